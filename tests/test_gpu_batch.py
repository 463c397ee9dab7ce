"""Batched persistent form: every replica's forces against the oracle, every
replica's trajectory against a single-system context with the same seed and
state (device noise; the Cartesian and the normal-mode propagator; Langevin,
Nose-Hoover chains with either coupling, or no thermostat), graph
determinism and the autotuner's state restore."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, KELVIN, MEV, aziz_config, per_bead_err, rel_err

pytestmark = pytest.mark.gpu


def dipole_replicas(n, seed=0):
    """Config-4-like replicas: N in {8..128}, P in {16,32,64}, T in [1,4] K,
    D=2, trap 3 meV + dipole pairs, all pairs; particles on a jittered 4-bohr
    lattice so the 1/r^4 dipole forces stay integrable at dt = 20 a.u."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        N = int(rng.choice(np.arange(8, 129, 8)))
        P = int(rng.choice([16, 32, 64]))
        T = rng.uniform(1.0, 4.0)
        kw = dict(natoms=N, nbeads=P, ndim=2, mass=4.0026 * AMU, temperature=T * KELVIN, dt=20.0,
                  gamma=1.0 / 2000.0, bosonic=1, pbc=0, size=60.0, ext_kind="harmonic", ext_omega=3.0 * MEV,
                  int_kind="dipole", int_strength=rng.uniform(0.5, 2.0), cutoff=-1.0, propagator="cartesian",
                  thermostat="langevin", nmthermostat=int(r % 2), fixcom=int(r % 3 == 0), seed=1000 + r)
        side = int(np.ceil(np.sqrt(N)))
        grid = np.array([(a, c) for a in range(side) for c in range(side)][:N], dtype=float) * 4.0
        centres = grid - grid.mean(axis=0) + rng.uniform(-0.5, 0.5, size=(N, 2))   # >= 3 bohr apart
        x = np.repeat(centres[None], P, axis=0) + rng.normal(0, 0.3, size=(P, N, 2))
        p = rng.normal(0, 0.5, size=(P, N, 2))
        out.append((kw, x, p))
    return out


def aziz_replicas(n):
    out = []
    for r in range(n):
        kw, x = aziz_config(int([32, 64, 48, 96][r % 4]), int([8, 16, 32, 8][r % 4]), 1.5 + 0.2 * r,
                            replicate=False, seed=40 + r)
        kw["seed"] = 77 + r
        p = np.random.default_rng(r).normal(0, 0.3, size=x.shape)
        out.append((kw, x, p))
    return out


def trap_replicas(n):
    """Replicas with the reference's other external potentials (double well,
    cosine) under harmonic pairs, uneven N and P."""
    rng = np.random.default_rng(21)
    out = []
    for r in range(n):
        N, P = int([12, 40, 24][r % 3]), int([8, 4, 16][r % 3])
        kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
                  gamma=1.0 / 2000.0, bosonic=int(r % 2 == 0), pbc=0, size=30.0,
                  ext_kind=["double_well", "cosine"][r % 2], ext_strength=1.0 * MEV, ext_location=1.9,
                  ext_amplitude=2.0 * MEV, ext_phase=0.3, int_kind="harmonic", int_omega=1.0 * MEV,
                  cutoff=-1.0, propagator="cartesian", thermostat="langevin", seed=500 + r)
        x = rng.uniform(-6.0, 6.0, size=(P, N, 3))
        p = rng.normal(0, 0.5, size=(P, N, 3))
        out.append((kw, x, p))
    return out


def nm_replicas(n):
    """Normal-mode Langevin replicas with bead counts that are not multiples
    of 16 and above 64 (the two forms of the batched normal-mode transform),
    uneven N, harmonic trap and pairs."""
    rng = np.random.default_rng(31)
    out = []
    for r in range(n):
        N, P = int([8, 20, 33, 12, 17, 9][r % 6]), int([5, 24, 40, 72, 96, 1][r % 6])
        kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
                  gamma=1.0 / 500.0, bosonic=int(r % 2 == 0), pbc=0, size=30.0, ext_kind="harmonic",
                  ext_omega=2.0 * MEV, int_kind="harmonic", int_omega=1.0 * MEV, cutoff=-1.0,
                  propagator="cartesian", thermostat="langevin", nmthermostat=1, fixcom=int(r % 3 != 1),
                  seed=900 + r)
        x = rng.uniform(-5.0, 5.0, size=(P, N, 3))
        p = rng.normal(0, 0.5, size=(P, N, 3))
        out.append((kw, x, p))
    return out


def nmprop_replicas(n):
    """Distinguishable replicas with the normal-mode propagator (the reference
    allows it only without exchange, params.cpp:146): normal-mode Langevin on
    the matrix cores (P <= 64) and in the VALU form (P = 72), Cartesian
    Langevin, no thermostat; with and without fixcom; uneven N, harmonic trap
    and pairs."""
    rng = np.random.default_rng(41)
    out = []
    for r in range(n):
        N, P = int([8, 20, 33, 12, 17, 9][r % 6]), int([8, 24, 64, 72, 16, 4][r % 6])
        th, nmth = [("langevin", 1), ("langevin", 1), ("langevin", 0), ("langevin", 1), ("none", 0),
                    ("langevin", 0)][r % 6]
        kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
                  gamma=1.0 / 500.0, bosonic=0, pbc=0, size=30.0, ext_kind="harmonic", ext_omega=2.0 * MEV,
                  int_kind="harmonic", int_omega=1.0 * MEV, cutoff=-1.0, propagator="normal_modes",
                  thermostat=th, nmthermostat=nmth, fixcom=int(r % 3 != 1), seed=700 + r)
        x = rng.uniform(-5.0, 5.0, size=(P, N, 3))
        p = rng.normal(0, 0.5, size=(P, N, 3))
        out.append((kw, x, p))
    return out


def nh_replicas(n):
    """Nose-Hoover chain replicas with Cartesian coupling (nose_hoover.cpp):
    one chain per bead, per particle and per degree of freedom, chain lengths
    2 to 5, with and without fixcom, bosonic and distinguishable, the
    Cartesian and (distinguishable) the normal-mode propagator; uneven N and
    P, harmonic trap and pairs."""
    rng = np.random.default_rng(51)
    out = []
    for r in range(n):
        N, P = int([8, 20, 13, 12, 17, 9][r % 6]), int([8, 24, 16, 32, 4, 12][r % 6])
        th = ["nose_hoover", "nose_hoover_np", "nose_hoover_np_dim"][r % 3]
        bosonic = int(r % 2 == 0)
        kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
                  gamma=1.0 / 500.0, bosonic=bosonic, pbc=0, size=30.0, ext_kind="harmonic",
                  ext_omega=2.0 * MEV, int_kind="harmonic", int_omega=1.0 * MEV, cutoff=-1.0,
                  propagator="normal_modes" if (not bosonic and r % 4 == 1) else "cartesian",
                  thermostat=th, nmthermostat=0, nchains=[4, 2, 5, 3, 4, 2][r % 6], fixcom=int(r % 3 != 1),
                  seed=800 + r)
        x = rng.uniform(-5.0, 5.0, size=(P, N, 3))
        p = rng.normal(0, 0.5, size=(P, N, 3))
        out.append((kw, x, p))
    return out


def nhnm_replicas(n):
    """nh_replicas with normal-mode coupling (the chains on the normal modes
    of the momenta, thermostat_coupling.cpp)."""
    out = []
    for kw, x, p in nh_replicas(n):
        kw = dict(kw)
        kw["nmthermostat"] = 1
        kw["seed"] += 100
        out.append((kw, x, p))
    return out


@pytest.mark.parametrize("which", ["dipole", "aziz", "traps", "nmprop", "nh", "nhnm"])
def test_batch_forces_match_oracle(which):
    reps = {"dipole": lambda: dipole_replicas(12, seed=3), "aziz": lambda: aziz_replicas(6),
            "traps": lambda: trap_replicas(6), "nmprop": lambda: nmprop_replicas(6),
            "nh": lambda: nh_replicas(6), "nhnm": lambda: nhnm_replicas(6)}[which]()
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    for r, (kw, x, p) in enumerate(reps):
        _, _, f = b.state(r)
        fo = O.update_forces(O.make_params(**kw), x)
        assert per_bead_err(f, fo) < 1e-10, r
    b.close()


def test_batch_aziz_ranges_periodic():
    """Aziz pairs in each range of the force under periodic boundaries with a
    cutoff (the batched walk in box-fractional coordinates): the hard-core
    range r/rm in (1e-7, 0.01) (repulsive term only, aziz.cpp:56-106), the
    damped range on both sides of 0.01 and just below D, just above D and the
    undamped tail (0.03 rm: the damping exponent below its clamp), each pair
    60 bohr from the next; a distinguishable and a bosonic replica in one
    batch, every pair's force against the oracle."""
    rm = 5.60738
    seps = np.array([0.005, 0.0099, 0.0101, 0.03, 0.3, 0.9, 1.2413, 1.2414, 1.6, 2.5]) * rm
    n, P = len(seps), 3
    rng = np.random.default_rng(23)
    x = np.zeros((P, 2 * n, 3))
    for k, s in enumerate(seps):
        c = np.array([60.0 * k + 5.0, 7.0, 3.0])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for b in range(P):
            x[b, 2 * k] = c
            x[b, 2 * k + 1] = c + s * u
    reps = []
    for bosonic in (0, 1):
        kw = dict(natoms=2 * n, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
                  gamma=1e-3, bosonic=bosonic, pbc=1, size=700.0, ext_kind="free", int_kind="aziz",
                  cutoff=250.0, propagator="cartesian", thermostat="langevin", seed=3 + bosonic)
        reps.append((kw, x))
    b = G.Batch([G.make_params(**kw) for kw, _ in reps])
    for r, (kw, xr) in enumerate(reps):
        b.set_state(r, xr, np.zeros_like(xr))
    b.forces()
    for r, (kw, xr) in enumerate(reps):
        _, _, f = b.state(r)
        fo = O.update_forces(O.make_params(**kw), xr)
        assert np.all(np.isfinite(f)), r
        for k in range(n):
            sl = slice(2 * k, 2 * k + 2)
            err = np.max(np.abs(f[:, sl] - fo[:, sl])) / np.max(np.abs(fo[:, sl]))
            assert err < 1e-10, (r, seps[k] / rm, err)
    b.close()


@pytest.mark.parametrize("which", ["dipole", "aziz", "nm", "nmprop", "nh", "nhnm"])
def test_batch_follows_single_system(which):
    reps = {"dipole": lambda: dipole_replicas(6, seed=5), "aziz": lambda: aziz_replicas(4),
            "nm": lambda: nm_replicas(6), "nmprop": lambda: nmprop_replicas(6),
            "nh": lambda: nh_replicas(6), "nhnm": lambda: nhnm_replicas(6)}[which]()
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    ctxs = []
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
        c = G.Context(G.make_params(**kw))
        c.set_state(x, p)
        c.forces()
        ctxs.append(c)
    b.forces()
    b.step(15)
    for r, c in enumerate(ctxs):
        c.step(15)
        xb, pb, fb = b.state(r)
        xc, pc, fc = c.state()
        assert rel_err(xb, xc) < 1e-10, r
        assert rel_err(pb, pc) < 1e-9, r
        assert per_bead_err(fb, fc) < 1e-9, r
        c.close()
    b.close()


@pytest.mark.parametrize("which", ["dipole", "aziz"])
def test_batch_graph_replays_are_deterministic(which):
    """Two batches of the same replicas agree bitwise after every replay
    (the Aziz case exercises the LDS force accumulation of the pair phase and
    the exchange's chunked exterior forces)."""
    reps = dipole_replicas(9, seed=8) if which == "dipole" else aziz_replicas(4)
    bs = [G.Batch([G.make_params(**kw) for kw, _, _ in reps]) for _ in range(2)]
    for b in bs:
        for r, (kw, x, p) in enumerate(reps):
            b.set_state(r, x, p)
        b.forces()
    for rep in range(3):
        for b in bs:
            b.step(4)
        for r in range(len(reps)):
            sa, sb = bs[0].state(r), bs[1].state(r)
            assert all(np.array_equal(u, v) for u, v in zip(sa, sb)), (rep, r)
    for b in bs:
        b.close()


def test_batch_autotune_restores_state():
    reps = dipole_replicas(5, seed=9)
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    before = [b.state(r) for r in range(len(reps))]
    threads, resident = b.autotune(trial_steps=2)
    assert threads in (256, 512, 1024) and resident >= 1
    after = [b.state(r) for r in range(len(reps))]
    for s0, s1 in zip(before, after):
        assert all(np.array_equal(u, v) for u, v in zip(s0, s1))
    b.step(3)
    b.close()


def test_batch_autotune_then_nm_steps_follow_single_system():
    """The autotuner's trial steps leave the state, the normal-mode replicas'
    physical forces and the Nose-Hoover chains included, as they were: the
    steps after it follow the single-system form."""
    reps = nmprop_replicas(3) + nh_replicas(3) + nhnm_replicas(2)
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    b.autotune(trial_steps=2)
    b.step(10)
    for r, (kw, x, p) in enumerate(reps):
        c = G.Context(G.make_params(**kw))
        c.set_state(x, p)
        c.forces()
        c.step(10)
        xb, pb, _ = b.state(r)
        xc, pc, _ = c.state()
        assert rel_err(xb, xc) < 1e-10, r
        assert rel_err(pb, pc) < 1e-9, r
        c.close()
    b.close()


def test_batch_rejects_mixed_dimensions():
    reps = dipole_replicas(2, seed=1)
    kw = dict(reps[1][0])
    kw["ndim"] = 3
    with pytest.raises(G.PimdbError) as ei:
        G.Batch([G.make_params(**reps[0][0]), G.make_params(**kw)])
    assert ei.value.code == G.E_ARG


def test_batch_state_reads_between_steps():
    """Reading the state mid-run (which completes the owed second O-step of the
    normal-mode replicas) leaves the trajectory unchanged up to rounding; the
    replicas include traps (forces that do not sum to zero) and fixcom."""
    reps = nm_replicas(6)
    bs = [G.Batch([G.make_params(**kw) for kw, _, _ in reps]) for _ in range(2)]
    for b in bs:
        for r, (kw, x, p) in enumerate(reps):
            b.set_state(r, x, p)
        b.forces()
    bs[0].step(10)
    bs[1].step(4)
    for r in range(len(reps)):
        bs[1].state(r)
    bs[1].step(3)
    bs[1].state(0)
    bs[1].step(3)
    for r in range(len(reps)):
        xa, pa, fa = bs[0].state(r)
        xb, pb, fb = bs[1].state(r)
        assert rel_err(xb, xa) < 1e-12, r
        assert rel_err(pb, pa) < 1e-11, r
        assert per_bead_err(fb, fa) < 1e-11, r
    for b in bs:
        b.close()


def test_batch_set_state_after_steps():
    """Replacing a replica's state after steps starts it afresh: it then
    follows the single-system form from the new state."""
    reps = nm_replicas(3)
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    b.step(5)
    kw, x, p = reps[1]
    b.set_state(1, x, p)
    b.forces()
    c = G.Context(G.make_params(**kw))
    c.set_state(x, p)
    c.forces()
    # the batch's step counter has advanced by 5: the single system skips 5 steps of noise keys
    c.step(5)
    c.set_state(x, p)
    c.forces()
    b.step(6)
    c.step(6)
    xb, pb, fb = b.state(1)
    xc, pc, fc = c.state()
    assert rel_err(xb, xc) < 1e-10
    assert rel_err(pb, pc) < 1e-9
    c.close()
    b.close()
