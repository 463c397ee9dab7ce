"""Energy / classical / bosonic observables computed on the GPU against the
oracle (summed over beads, atomic units) and against the reference's
simulation.out rows."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import KELVIN, aziz_config, case_params, load_fixture

pytestmark = pytest.mark.gpu

NAMES = ["kinetic", "potential", "ext_pot", "int_pot", "virial", "temperature", "cl_kinetic",
         "cl_spring", "prob_dist", "prob_all", "nh_energy"]


def compare(kw, x, p, rtol=1e-9):
    sim = O.Sim(O.make_params(**kw))
    sim.set_state(x, p)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    og = ctx.observables()
    oo = sim.observables()
    ctx.close()
    scale = np.maximum(np.abs(oo), 1e-300)
    for j, n in enumerate(NAMES):
        if n in ("prob_dist", "prob_all"):
            assert abs(og[j] - oo[j]) <= 1e-10 + rtol * abs(oo[j]), (n, og[j], oo[j])
        else:
            assert abs(og[j] - oo[j]) <= rtol * max(scale[j], 1e-12), (n, og[j], oo[j])


@pytest.mark.parametrize("replicate", [True, False])
def test_observables_aziz(replicate):
    kw, x = aziz_config(96, 16, 1.5, replicate=replicate, seed=31)
    p = np.random.default_rng(1).normal(0, 0.3, size=x.shape)
    compare(kw, x, p)


@pytest.mark.parametrize("kind", ["aziz", "dipole", "harmonic"])
def test_observables_pairs_in_a_trap(kind):
    """int_pot and the pair virial reported separately (both potentials
    non-free): Aziz under PBC with a cutoff, dipole and harmonic pairs without."""
    if kind == "aziz":
        kw, x = aziz_config(64, 8, 1.5, replicate=False, seed=41)
        kw.update(ext_kind="harmonic", ext_omega=3.0e-3 * 0.036749326)
    else:
        rng = np.random.default_rng(42)
        N, P = 24, 8
        grid = np.array([(a, b, c) for a in range(3) for b in range(3) for c in range(3)][:N], float) * 4.0
        x = np.stack([grid - grid.mean(0) + rng.normal(0, 0.3, size=grid.shape) for _ in range(P)])
        kw = dict(natoms=N, nbeads=P, ndim=3, mass=7296.3, temperature=2.0 * KELVIN, dt=20.0, gamma=5e-4,
                  bosonic=1, pbc=0, size=40.0, ext_kind="harmonic", ext_omega=1.1e-4, int_kind=kind,
                  int_omega=3.7e-5, int_strength=1.3, cutoff=-1.0, propagator="cartesian",
                  thermostat="langevin", seed=9)
    p = np.random.default_rng(2).normal(0, 0.3, size=x.shape)
    compare(kw, x, p)


@pytest.mark.parametrize("ext", ["double_well", "cosine"])
def test_observables_external_traps(ext):
    """Energy, virial and their ext/int split with the reference's double-well
    and cosine external potentials (harmonic pairs)."""
    rng = np.random.default_rng(6)
    N, P = 20, 8
    x = rng.uniform(-5.0, 5.0, size=(P, N, 3))
    p = rng.normal(0, 0.4, size=x.shape)
    kw = dict(natoms=N, nbeads=P, ndim=3, mass=7296.3, temperature=2.0 * KELVIN, dt=20.0, gamma=5e-4,
              bosonic=1, pbc=0, size=25.0, ext_kind=ext, ext_strength=3.7e-5, ext_location=1.9,
              ext_amplitude=7.3e-5, ext_phase=0.3, int_kind="harmonic", int_omega=3.7e-5, cutoff=-1.0,
              propagator="cartesian", thermostat="langevin", seed=9)
    compare(kw, x, p)


@pytest.mark.parametrize("case", ["bosonic_quadratic_harmonic", "dist_harmonic",
                                  "bosonic_quadratic_harmonic_nmthermostat_dynamics", "bosonic_factorial_harmonic",
                                  "bosonic_quadratic_harmonic_nh_dynamics", "bosonic_quadratic_harmonic_nh_np_dynamics",
                                  "bosonic_quadratic_harmonic_nh_np_dim_dynamics"])
def test_observables_reference_rows(case):
    """Steps on the GPU (reference noise) then its observables against the
    stored simulation.out rows (printed with 9 significant digits)."""
    kw, _ = case_params(case)
    fx = load_fixture(case)
    header = [str(h) for h in fx["sim_out_header"]]
    sim = O.Sim(O.make_params(**kw))
    sim.init_random()
    x, p, _ = sim.state()
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    done = 0
    for row in fx["sim_out"][:4]:
        n = int(row[0]) + 1 - done
        noise = np.stack([sim.draw_noise() for _ in range(n)])
        ctx.step(n, noise)
        done += n
        ob = ctx.observables() / KELVIN
        for j, name in enumerate(header[1:]):
            if name in ("ext_pot", "int_pot"):
                continue
            ref = row[1 + j]
            assert abs(ob[NAMES.index(name)] - ref) <= 2e-8 * abs(ref) + 1e-8, (name, ob, ref)
    ctx.close()


@pytest.mark.parametrize("N,P,replicate", [(8, 4, False), (96, 16, True), (256, 8, False)])
def test_exchange_estimators(N, P, replicate):
    """primEstimator / getDistinctProbability / getLongestProbability
    (quadratic_bosonic_exchange.cpp:222-279) from the GPU entry point against
    the oracle's restatement on the same positions."""
    kw, x = aziz_config(N, P, 1.5, replicate=replicate, seed=5)
    prm = O.make_params(**kw)
    d = O.derived(prm)
    E = O.cycle_energies(x[0], x[-1], d["k"], True, kw["size"])
    ex = O.exchange(x[0], x[1 % P], x[P - 2], x[-1], d["beta_t"], d["k"], True, kw["size"])
    L = O.lib()
    prim = L.orc_prim_estimator(N, P, d["beta_t"], O.ptr(E), O.ptr(ex["V"]))
    pd = L.orc_distinct_probability(N, d["beta_t"], O.ptr(E), O.ptr(ex["V"]))
    pl = L.orc_longest_probability(N, d["beta_t"], O.ptr(E), O.ptr(ex["V"]))
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, np.zeros_like(x))
    g_prim, g_pd, g_pl = ctx.exchange_estimators()
    ctx.close()
    assert abs(g_prim - prim) <= 1e-10 * max(abs(prim), 1e-300), (g_prim, prim)
    assert abs(g_pd - pd) <= 1e-10 * max(abs(pd), 1e-300) + 1e-300, (g_pd, pd)
    assert abs(g_pl - pl) <= 1e-10 * max(abs(pl), 1e-300) + 1e-300, (g_pl, pl)


def test_exchange_estimators_need_bosons():
    kw, x = aziz_config(16, 4, 1.5, seed=5)
    kw["bosonic"] = 0
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, np.zeros_like(x))
    with pytest.raises(G.PimdbError):
        ctx.exchange_estimators()
    ctx.close()
