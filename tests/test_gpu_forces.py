"""Forces on every bead (exchange + interior springs + trap + pair
interactions with cutoff and minimum image) against the oracle's
updateForces restatement, norm-wise per bead <= 1e-10."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, KELVIN, MEV, aziz_config, per_bead_err

pytestmark = pytest.mark.gpu


def gpu_forces(kw, x):
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x)
    ctx.forces()
    _, _, f = ctx.state()
    ctx.close()
    return f


CASES = {
    "aziz_pbc_bosonic": dict(N=40, P=8, D=3, int_kind="aziz", pbc=1, L=25.0, cutoff=100.0, bosonic=1, ext="free"),
    "aziz_pbc_dist": dict(N=40, P=5, D=3, int_kind="aziz", pbc=1, L=25.0, cutoff=100.0, bosonic=0, ext="free"),
    "aziz_open_cut": dict(N=50, P=4, D=3, int_kind="aziz", pbc=0, L=25.0, cutoff=9.0, bosonic=1, ext="harmonic"),
    "dipole_2d": dict(N=64, P=16, D=2, int_kind="dipole", pbc=0, L=40.0, cutoff=-1.0, bosonic=1, ext="harmonic"),
    "harmonic_pairs": dict(N=30, P=8, D=3, int_kind="harmonic", pbc=0, L=40.0, cutoff=-1.0, bosonic=1, ext="harmonic"),
    "harmonic_pbc_1d": dict(N=20, P=6, D=1, int_kind="harmonic", pbc=1, L=15.0, cutoff=100.0, bosonic=1, ext="harmonic"),
    "free_trap": dict(N=33, P=7, D=3, int_kind="free", pbc=0, L=40.0, cutoff=-1.0, bosonic=1, ext="harmonic"),
    "one_bead": dict(N=25, P=1, D=3, int_kind="aziz", pbc=1, L=20.0, cutoff=100.0, bosonic=1, ext="free"),
    "two_beads": dict(N=25, P=2, D=3, int_kind="dipole", pbc=1, L=20.0, cutoff=100.0, bosonic=1, ext="free"),
    "double_well_harmonic": dict(N=30, P=8, D=3, int_kind="harmonic", pbc=0, L=30.0, cutoff=-1.0, bosonic=1,
                                 ext="double_well"),
    "double_well_2d_dist": dict(N=24, P=6, D=2, int_kind="free", pbc=0, L=30.0, cutoff=-1.0, bosonic=0,
                                ext="double_well"),
    "cosine_dipole": dict(N=40, P=8, D=3, int_kind="dipole", pbc=1, L=30.0, cutoff=100.0, bosonic=1, ext="cosine"),
    "cosine_1d": dict(N=16, P=4, D=1, int_kind="harmonic", pbc=1, L=12.0, cutoff=100.0, bosonic=1, ext="cosine"),
}

# external double well: strength 1 meV at 1 angstrom; cosine: amplitude 2 meV, phase 0.3
EXT = dict(ext_strength=1.0 * MEV, ext_location=1.8897261, ext_amplitude=2.0 * MEV, ext_phase=0.3)


@pytest.mark.parametrize("name", sorted(CASES))
def test_forces_match_oracle(name):
    c = CASES[name]
    rng = np.random.default_rng(abs(hash(name)) % 2 ** 31)
    L = c["L"]
    x = rng.uniform(-0.5 * L, 0.5 * L, size=(c["P"], c["N"], c["D"]))
    kw = dict(natoms=c["N"], nbeads=c["P"], ndim=c["D"], mass=4.0026 * AMU, temperature=2.0 * KELVIN,
              dt=20.0, gamma=1e-3, bosonic=c["bosonic"], pbc=c["pbc"], size=L,
              ext_kind=c["ext"], ext_omega=3.0 * MEV, int_kind=c["int_kind"], int_omega=1.0 * MEV,
              int_strength=0.8, cutoff=c["cutoff"], propagator="cartesian", thermostat="langevin", seed=3, **EXT)
    f = gpu_forces(kw, x)
    fo = O.update_forces(O.make_params(**kw), x)
    assert per_bead_err(f, fo) < 1e-10


@pytest.mark.parametrize("replicate", [True, False])
def test_forces_headline_aziz(replicate):
    """N=256, P=64, He-4 Aziz, PBC cutoff L/2, T=1.5 K."""
    kw, x = aziz_config(256, 64, 1.5, replicate=replicate, seed=2024)
    f = gpu_forces(kw, x)
    fo = O.update_forces(O.make_params(**kw), x)
    assert per_bead_err(f, fo) < 1e-10


def test_forces_translation_invariance_without_trap():
    """Pair + spring + exchange forces sum to zero over all beads (Newton's
    third law and translation invariance of V_B), size-independent check."""
    kw, x = aziz_config(128, 32, 1.5, replicate=False, seed=99)
    f = gpu_forces(kw, x)
    tot = f.sum(axis=(0, 1))
    assert np.max(np.abs(tot)) < 1e-10 * np.max(np.abs(f)) * f.shape[0] * f.shape[1]


@pytest.mark.parametrize("bosonic", [0, 1])
def test_forces_aziz_ranges(bosonic):
    """Aziz pairs placed in each range of the force: the hard-core range
    r/rm in (1e-7, 0.01) (repulsive term only, aziz.cpp), the damped range
    r/rm < D = 1.241314 on both sides of 0.01 and just below D, just above D,
    and the undamped tail; each pair far from the others."""
    rm = 5.60738
    seps = np.array([0.005, 0.0099, 0.0101, 0.3, 0.9, 1.2413, 1.2414, 1.6, 2.5]) * rm
    n = len(seps)
    P = 3
    rng = np.random.default_rng(17)
    x = np.zeros((P, 2 * n, 3))
    for k, s in enumerate(seps):
        c = np.array([60.0 * k, 0.0, 0.0])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for b in range(P):
            x[b, 2 * k] = c
            x[b, 2 * k + 1] = c + s * u
    kw = dict(natoms=2 * n, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0, gamma=1e-3,
              bosonic=bosonic, pbc=0, size=100.0, ext_kind="free", int_kind="aziz", cutoff=-1.0,
              propagator="cartesian", thermostat="langevin", seed=3)
    f = gpu_forces(kw, x)
    fo = O.update_forces(O.make_params(**kw), x)
    # pair by pair: each pair's force relative to its own scale
    for k in range(n):
        sl = slice(2 * k, 2 * k + 2)
        err = np.max(np.abs(f[:, sl] - fo[:, sl])) / np.max(np.abs(fo[:, sl]))
        assert err < 1e-10, (seps[k] / rm, err)
