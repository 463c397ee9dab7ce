"""MD steps on the GPU: with the reference's RanMars noise fed from the host,
trajectories follow the oracle (and the reference's stored frames); with
device noise, graph replays equal eager launches bit for bit."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import (ANGSTROM, FORCE_EV_A, VEL_A_PS, aziz_config, case_params, load_fixture,
                     per_bead_err, rel_err)

pytestmark = pytest.mark.gpu


def run_pair(kw, nsteps, x=None, p=None, init_random=False, forces_first=False):
    """Oracle and GPU from the same state with the same RanMars noise."""
    sim = O.Sim(O.make_params(**kw))
    if init_random:
        sim.init_random()
    else:
        sim.set_state(x, p)
    if forces_first:
        sim.forces()
    x0, p0, _ = sim.state()
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x0, p0)
    if forces_first:
        ctx.forces()
    noise = np.stack([sim.draw_noise() for _ in range(nsteps)])
    for s in range(nsteps):
        sim.step(noise[s])
    ctx.step(nsteps, noise)
    g = ctx.state()
    ctx.close()
    return g, sim.state()


CASES = ["bosonic_quadratic_harmonic_dynamics", "bosonic_quadratic_harmonic_nmthermostat_dynamics",
         "dist_harmonic_dynamics", "dist_harmonic_nm_propagation_dynamics", "bosonic_factorial_harmonic_dynamics",
         "bosonic_quadratic_harmonic_nh_dynamics", "bosonic_quadratic_harmonic_nh_np_dynamics",
         "bosonic_quadratic_harmonic_nh_np_dim_dynamics"]


@pytest.mark.parametrize("case", CASES)
def test_trajectory_follows_oracle(case):
    kw, _ = case_params(case)
    (xg, pg, fg), (xo, po, fo) = run_pair(kw, 300, init_random=True)
    assert rel_err(xg, xo) < 1e-10
    assert rel_err(pg, po) < 1e-10
    assert per_bead_err(fg, fo) < 1e-10


@pytest.mark.parametrize("case", CASES)
def test_trajectory_reaches_reference_frame(case):
    """sfreq steps from the reference's initial state land on its next stored
    frame (positions, velocities, forces as printed, 13 digits)."""
    kw, sfreq = case_params(case)
    fx = load_fixture(case)
    (xg, pg, fg), _ = run_pair(kw, sfreq, init_random=True)
    assert rel_err(xg / ANGSTROM, fx["positions"][1]) < 1e-9
    assert rel_err(pg / kw["mass"] / VEL_A_PS, fx["velocities"][1]) < 1e-9
    assert per_bead_err(fg / FORCE_EV_A, fx["forces"][1]) < 1e-9


def test_aziz_fixcom_trajectory():
    """He-4 Aziz, PBC, NM thermostat, fixcom: 60 steps with host noise."""
    kw, x = aziz_config(64, 16, 1.5, replicate=False, seed=5)
    rng = np.random.default_rng(6)
    p = rng.normal(0, 0.3, size=x.shape)
    (xg, pg, fg), (xo, po, fo) = run_pair(kw, 60, x=x, p=p, forces_first=True)
    assert rel_err(xg, xo) < 1e-10
    assert rel_err(pg, po) < 1e-9
    assert per_bead_err(fg, fo) < 1e-9
    # zeroMomentum leaves no centre-of-mass momentum
    assert np.max(np.abs(pg.sum(axis=(0, 1)))) < 1e-12 * np.max(np.abs(pg)) * pg.shape[0] * pg.shape[1]


def test_graph_replay_equals_eager():
    """Device (Philox) noise: a captured step replayed N times equals N eager
    steps exactly; replays interleaved with host work and syncs."""
    kw, x = aziz_config(48, 8, 1.5, replicate=False, seed=8)
    rng = np.random.default_rng(9)
    p = rng.normal(0, 0.3, size=x.shape)
    a = G.Context(G.make_params(**kw))
    b = G.Context(G.make_params(**kw))
    for c in (a, b):
        c.set_state(x, p)
        c.forces()
    b.graph_capture()
    for rep in range(4):
        a.step(5)
        b.step(5)
        xa, pa, fa = a.state()
        xb, pb, fb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(pa, pb) and np.array_equal(fa, fb), rep
    a.close()
    b.close()


def test_device_noise_thermalizes():
    """Free distinguishable particles with the Langevin thermostat: the
    ring-polymer kinetic temperature relaxes to P*T (IPI convention)."""
    N, P, D = 64, 8, 3
    T = 1e-4
    kw = dict(natoms=N, nbeads=P, ndim=D, mass=50.0, temperature=T, dt=5.0, gamma=1.0 / 50.0,
              bosonic=0, pbc=0, size=50.0, ext_kind="harmonic", ext_omega=0.002, int_kind="free",
              propagator="cartesian", thermostat="langevin", nmthermostat=0, fixcom=0, seed=77)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(np.zeros((P, N, D)), np.zeros((P, N, D)))
    ctx.step(2000)
    temps = []
    for _ in range(40):
        ctx.step(25)
        temps.append(ctx.observables()[5])   # quantum temperature estimator (classical / P)
    ctx.close()
    assert abs(np.mean(temps) / T - 1.0) < 0.05


@pytest.mark.parametrize("kind", ["nose_hoover", "nose_hoover_np", "nose_hoover_np_dim"])
@pytest.mark.parametrize("fixcom", [0, 1])
def test_nose_hoover_normal_mode_coupling_follows_oracle(kind, fixcom):
    """Nose-Hoover chains on the normal modes (nmthermostat; no reference
    fixture covers it): 200 steps against the oracle, bosonic and
    distinguishable rings, with and without zeroMomentum."""
    for bosonic in (1, 0):
        kw, _ = case_params("bosonic_quadratic_harmonic_nh_dynamics")
        kw.update(thermostat=kind, nmthermostat=1, fixcom=fixcom, bosonic=bosonic, nchains=3, nbeads=6)
        (xg, pg, fg), (xo, po, fo) = run_pair(kw, 200, init_random=True)
        assert rel_err(xg, xo) < 1e-10, bosonic
        assert rel_err(pg, po) < 1e-10, bosonic
        assert per_bead_err(fg, fo) < 1e-10, bosonic
