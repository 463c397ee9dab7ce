"""The factorial bosonic exchange on the GPU (the reference's compile-time
alternative, factorial_bosonic_exchange.cpp) against the oracle's restatement:
forces on every bead, V_eff and the kinetic term, on random rings with and
without PBC, and its argument checks."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, KELVIN, MEV, per_bead_err

pytestmark = pytest.mark.gpu


def ring_kw(N, P, D, pbc, seed):
    rng = np.random.default_rng(seed)
    L = 12.0
    centres = rng.uniform(-0.3 * L, 0.3 * L, size=(N, D))
    x = centres[None] + rng.normal(0, 0.8, size=(P, N, D))
    kw = dict(natoms=N, nbeads=P, ndim=D, mass=4.0026 * AMU, temperature=1.5 * KELVIN, dt=20.0, gamma=1e-3,
              bosonic=1, pbc=pbc, size=L, ext_kind="harmonic", ext_omega=2.0 * MEV, int_kind="free",
              cutoff=0.0, propagator="cartesian", thermostat="langevin", seed=7, exchange_algo="factorial")
    return kw, x


CASES = [(2, 4, 3, 0), (3, 8, 3, 1), (5, 6, 2, 0), (6, 5, 1, 1), (7, 4, 3, 1), (8, 8, 3, 0)]


@pytest.mark.parametrize("N,P,D,pbc", CASES)
def test_factorial_forces_match_oracle(N, P, D, pbc):
    kw, x = ring_kw(N, P, D, pbc, seed=N * 100 + P)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x)
    ctx.forces()
    _, _, f = ctx.state()
    fo = O.update_forces(O.make_params(**kw), x)
    assert per_bead_err(f, fo) < 1e-10
    # the adjoint entry point gives the exchange forces alone
    f1, fP = ctx.exchange_adjoint()
    d = O.derived(O.make_params(**kw))
    o = O.factorial_exchange(N, D, d["beta_t"], d["k"], pbc, kw["size"], x[0], x[P - 1], x[1 % P],
                             x[(P - 2) % P])
    assert np.max(np.abs(f1 - o["f_first"])) <= 1e-10 * np.max(np.abs(o["f_first"]))
    assert np.max(np.abs(fP - o["f_last"])) <= 1e-10 * np.max(np.abs(o["f_last"]))
    vb = ctx.exchange_forward(prob=False, E=False, V=False)["VB"]
    assert abs(vb - o["vb"]) <= 1e-11 * abs(o["vb"])
    prim, pd, pl = ctx.exchange_estimators()
    assert abs(prim - o["prim"] / P) <= 1e-11 * abs(o["prim"] / P)
    assert pd == 0.0 and pl == 0.0
    ctx.close()


def test_factorial_rejects_what_it_does_not_define():
    kw, x = ring_kw(4, 4, 3, 0, seed=3)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x)
    with pytest.raises(G.PimdbError) as ei:
        ctx.exchange_forward(prob=False, E=False, V=True)
    assert ei.value.code == G.E_ARG
    ctx.close()
    kw11, _ = ring_kw(11, 4, 3, 0, seed=4)
    with pytest.raises(G.PimdbError) as ei:
        G.Context(G.make_params(**kw11))
    assert ei.value.code == G.E_ARG
    with pytest.raises(G.PimdbError) as ei:
        G.Batch([G.make_params(**kw)])
    assert ei.value.code == G.E_ARG
