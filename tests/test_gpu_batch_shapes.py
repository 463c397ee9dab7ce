"""The batched persistent kernel at the shapes it is measured on.

* Config 3 (N=256, P=64, Aziz, PBC at 0.02186 A^-3, cutoff L/2, beads not
  replicated) beside replicas of N = 129, 192, 255, 257 and 320, so that the
  two-row tile pass, ragged last tiles and the spatial tile order all run,
  with every replica's forces checked against the oracle per bead.
* The same replicas planned for two workgroups per compute unit (512
  threads, half the LDS each): the recursions then run one after the other.
* A distinguishable N=448 replica (its slice forces do not fit the LDS: the
  pair walk accumulates the columns in registers) and a bosonic N=780
  replica (the exterior forces get fewer chunk slots than waves, and the
  recursions run one after the other).
* A ring whose two-particle cycles dominate (one-particle cycles suppressed
  by 2^-430 each), beside a weakly paired one.
* 100 device-noise steps of the benchmark's own initial state through the
  batch against the single-system context with the same seed.

Reference: Simulation::updatePhysicalForces (src/simulation.cpp:428-455),
BosonicExchange (quadratic_bosonic_exchange.cpp:34-215), Simulation::run
(simulation.cpp:246-259).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, KELVIN, MEV, aziz_config, per_bead_err, rel_err

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the benchmark's workload generator)

pytestmark = pytest.mark.gpu


def headline_replicas():
    out = []
    for r, (N, P) in enumerate([(256, 64), (129, 8), (192, 8), (255, 8), (257, 8), (320, 8)]):
        kw, x = aziz_config(N, P, 1.5, replicate=False, seed=300 + r)
        kw["seed"] = 11 + r
        p = np.random.default_rng(r).normal(0, 0.3, size=x.shape)
        out.append((kw, x, p))
    return out


def check_forces(reps, threads=0, resident=0):
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    if threads or resident:
        b.configure(threads, resident)
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    errs = []
    for r, (kw, x, p) in enumerate(reps):
        _, _, f = b.state(r)
        fo = O.update_forces(O.make_params(**kw), x)
        errs.append(per_bead_err(f, fo))
    b.close()
    return errs


def test_headline_shapes_forces_match_oracle():
    errs = check_forces(headline_replicas())
    print("per-replica force errors", errs)
    assert max(errs) < 1e-10, errs


def test_headline_shapes_two_workgroups_per_cu():
    """512 threads, LDS planned for two workgroups per CU: the recursions run
    one after the other on all 8 waves."""
    cus = G.device_cus()
    errs = check_forces(headline_replicas(), threads=512, resident=2 * cus)
    print("per-replica force errors", errs)
    assert max(errs) < 1e-10, errs


def test_large_replicas_forces_match_oracle():
    """N=448 distinguishable (column forces in registers, no LDS slices) and
    N=780 bosonic (fewer exterior-force chunk slots than waves, recursions one
    after the other) in one batch."""
    reps = []
    for r, (N, bos) in enumerate([(448, 0), (780, 1)]):
        kw, x = aziz_config(N, 4, 1.5, replicate=False, seed=500 + r)
        kw["bosonic"] = bos
        p = np.random.default_rng(r).normal(0, 0.3, size=x.shape)
        reps.append((kw, x, p))
    errs = check_forces(reps)
    print("per-replica force errors", errs)
    assert max(errs) < 1e-10, errs


def paired_ring(N=128, P=8, sep=40.0, seed=3):
    """A ring in which every particle 2i shares a two-particle cycle with
    2i+1 (bead P of each sits at bead 1 of the other, `sep` bohr from its own
    bead 1).  At sep = 40 one-particle cycles cost about 430 in the base-2
    exponent against 0.3 for the pairs: the in-block min-plus shifts span
    thousands of units and every probability but the pairings underflows."""
    kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
              gamma=1.0 / 2000.0, bosonic=1, pbc=0, size=60.0, ext_kind="harmonic", ext_omega=0.01 * MEV,
              int_kind="free", cutoff=-1.0, propagator="cartesian", thermostat="langevin", seed=5)
    rng = np.random.default_rng(seed)
    x = np.zeros((P, N, 3))
    for i in range(N // 2):
        A = np.array([i * 100.0, 0.0, 0.0])
        B = A + np.array([0.0, sep, 0.0])
        f = (np.arange(P) / (P - 1))[:, None]
        x[:, 2 * i] = A + f * (B - A)
        x[:, 2 * i + 1] = B + f * (A - B)
    x += rng.normal(0, 0.3, x.shape)
    x -= x.mean(axis=(0, 1))
    p = rng.normal(0, 0.3, size=x.shape)
    return kw, x, p


def test_paired_rings_forces_match_oracle():
    """Forces of the strongly paired ring (sep 40 bohr) and of a weakly paired
    one (N=100, sep 1 bohr) against the oracle per bead."""
    kw, x, p = paired_ring()
    d = O.derived(O.make_params(**kw))
    o = O.exchange(x[0], x[1], x[-2], x[-1], d["beta_t"], d["k"], False, 0.0)
    assert o["prob"][1, 0] > 0.999   # particle 0 closes its cycle through particle 1
    kw2, x2, p2 = paired_ring(N=100, sep=1.0, seed=4)
    errs = check_forces([(kw, x, p), (kw2, x2, p2)])
    print("per-replica force errors", errs)
    assert max(errs) < 1e-10, errs


@pytest.mark.parametrize("shape", ["1024x1", "512x2"])
def test_benchmark_state_100_steps_follow_single_system(shape):
    """The benchmark's initial state (bench.workload: seeded uniform positions
    with a 2.2 A rejection, replicated over beads, Maxwell-Boltzmann momenta),
    100 device-noise steps through the batch (normal-mode Langevin on the
    matrix cores with the fused O-steps, fixcom) and through the single-system
    context with the same seed."""
    kw, x, p = bench.workload()
    b = G.Batch([G.make_params(**kw)])
    if shape == "512x2":
        b.configure(512, 2 * G.device_cus())
    b.set_state(0, x, p)
    b.forces()
    c = G.Context(G.make_params(**kw))
    c.set_state(x, p)
    c.forces()
    b.step(100)
    c.step(100)
    xb, pb, fb = b.state(0)
    xc, pc, fc = c.state()
    ex, ep, ef = rel_err(xb, xc), rel_err(pb, pc), per_bead_err(fb, fc)
    print("100-step differences x %.2e p %.2e f %.2e" % (ex, ep, ef))
    assert ex < 1e-10 and ep < 1e-9 and ef < 1e-9
    # the forces after 100 steps against the oracle on the batch's positions
    fo = O.update_forces(O.make_params(**kw), xb)
    assert per_bead_err(fb, fo) < 1e-10
    b.close()
    c.close()
