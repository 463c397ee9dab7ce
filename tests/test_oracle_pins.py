"""Pins for the oracle functions that no reference fixture reaches.

1. Topology probabilities (quadratic_bosonic_exchange.cpp:222-240).  With
   Z = exp(-beta V_B) = sum over compositions c of 0..N-1 into consecutive
   blocks [s..e] of prod (1/(e+1)) exp(-beta E(s,e)) (the expansion of the
   forward recursion, :73-99), the distinct probability is the all-singleton
   composition's share, w(singletons) / Z, and the longest-cycle probability
   is N w(one block) / Z.  Checked against that enumeration (N <= 10, cycle
   energies from their closed form, computed here from the positions) and
   against closed forms: N = 1, N = 2, coincident beads (1/N!, 1), far-apart
   particles (1, 0).  A dropped ln N!, E(N,1) in place of E(N,N) or V[N-1] in
   place of V[N] fails these.
2. The pair terms of the energy observable (src/observables/energy.cpp:56-111):
   per bead, int_pot = sum_{i<j in cutoff} V(d_ij) with d_ij = x_i - x_j
   (minimum image) and virial = -sum_i x_i . F_ext,i - sum_{i<j} x_i . F_{i<-j}
   (only particle i's coordinate: quirk Q8), then potential / P, virial 0.5/P,
   ext_pot and int_pot / P when both potentials are non-free.  Worked here for
   2 and 3 particles with harmonic, dipole and Aziz pairs (with PBC and a pair
   beyond the cutoff); a virial that also adds x_j . F_{j<-i} fails it.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O
from helpers import AMU, ANGSTROM, KELVIN, MEV


def closed_E(x1, xP, k, u, v):
    """E of the cycle through particles u..v (u <= v): k/2 (sum_{j=u}^{v-1}
    |x1[j+1] - xP[j]|^2 + |x1[u] - xP[v]|^2), no periodic images."""
    s = sum(float(np.sum((x1[j + 1] - xP[j]) ** 2)) for j in range(u, v))
    return 0.5 * k * (s + float(np.sum((x1[u] - xP[v]) ** 2)))


def compositions(N):
    """Every composition of 0..N-1 into consecutive blocks, as (s, e) lists."""
    for cuts in itertools.product([0, 1], repeat=N - 1):
        blocks, s = [], 0
        for j, c in enumerate(cuts):
            if c:
                blocks.append((s, j))
                s = j + 1
        blocks.append((s, N - 1))
        yield blocks


def log_weight(blocks, x1, xP, k, beta):
    return sum(-math.log(e + 1) - beta * closed_E(x1, xP, k, s, e) for s, e in blocks)


def enumerated_probs(x1, xP, k, beta):
    N = x1.shape[0]
    logs = [log_weight(b, x1, xP, k, beta) for b in compositions(N)]
    m = max(logs)
    logZ = m + math.log(sum(math.exp(l - m) for l in logs))
    singles = log_weight([(i, i) for i in range(N)], x1, xP, k, beta)
    one = log_weight([(0, N - 1)], x1, xP, k, beta)
    return math.exp(singles - logZ), N * math.exp(one - logZ)


def oracle_probs(x1, xP, k, beta):
    N = x1.shape[0]
    ex = O.exchange(x1, x1, xP, xP, beta, k)
    L = O.lib()
    pd = L.orc_distinct_probability(N, beta, O.ptr(ex["E"]), O.ptr(ex["V"]))
    pa = L.orc_longest_probability(N, beta, O.ptr(ex["E"]), O.ptr(ex["V"]))
    return pd, pa


@pytest.mark.parametrize("N,seed", [(2, 1), (3, 2), (5, 3), (7, 4), (10, 5)])
def test_topology_probabilities_match_composition_enumeration(N, seed):
    rng = np.random.default_rng(seed)
    k, beta = 1.0, 1.0
    x1 = rng.normal(0.0, 0.7, size=(N, 3))
    xP = x1 + rng.normal(0.0, 0.5, size=(N, 3))
    pd, pa = oracle_probs(x1, xP, k, beta)
    ed, ea = enumerated_probs(x1, xP, k, beta)
    assert 1e-6 < ed < 1.0 and 1e-6 < ea < 1.0   # both probabilities far from 0 and 1
    assert abs(pd - ed) <= 1e-12 * ed, (pd, ed)
    assert abs(pa - ea) <= 1e-12 * ea, (pa, ea)


def test_topology_probabilities_closed_forms():
    rng = np.random.default_rng(11)
    k, beta = 0.8, 1.3
    # N = 1: the only configuration is both
    x1 = rng.normal(size=(1, 3))
    xP = x1 + rng.normal(size=(1, 3))
    pd, pa = oracle_probs(x1, xP, k, beta)
    assert abs(pd - 1.0) < 1e-14 and abs(pa - 1.0) < 1e-14
    # N = 2: Z = (e^{-b(E00+E11)} + e^{-b E01}) / 2, so the distinct share is
    # a / (a + b) and the longest 2 (b / 2) / Z = 2 b / (a + b)
    x1 = rng.normal(0.0, 0.4, size=(2, 3))
    xP = x1 + rng.normal(0.0, 0.4, size=(2, 3))
    e00, e11, e01 = (closed_E(x1, xP, k, u, v) for u, v in ((0, 0), (1, 1), (0, 1)))
    a, b = math.exp(-beta * (e00 + e11)), math.exp(-beta * e01)
    pd, pa = oracle_probs(x1, xP, k, beta)
    assert 1e-3 < b / (a + b) < 0.999
    assert abs(pd - a / (a + b)) < 1e-13 and abs(pa - 2.0 * b / (a + b)) < 1e-13
    # coincident beads: every E is zero, V[m] = 0, so 1/N! and 1
    for N in (3, 6):
        x = np.tile(rng.normal(size=(1, 3)), (N, 1))
        pd, pa = oracle_probs(x, x.copy(), k, beta)
        assert abs(pd - 1.0 / math.factorial(N)) < 1e-14 / math.factorial(N)
        assert abs(pa - 1.0) < 1e-13
    # particles far apart (each ring closed on itself): distinct -> 1, longest -> 0
    N = 5
    x1 = np.arange(N)[:, None] * np.array([[40.0, 0.0, 0.0]])
    xP = x1 + rng.normal(0.0, 0.1, size=(N, 3))
    pd, pa = oracle_probs(x1, xP, 1.0, 1.0)
    assert abs(pd - 1.0) < 1e-12 and pa < 1e-300


def pair_case(kind, n, pbc, seed):
    rng = np.random.default_rng(seed)
    P = 2
    kw = dict(natoms=n, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=20.0,
              gamma=1e-3, bosonic=0, pbc=int(pbc), ext_kind="harmonic", ext_omega=3.0 * MEV,
              int_kind=kind, int_omega=1.0 * MEV, int_strength=1.7, propagator="cartesian",
              thermostat="langevin", seed=3)
    if kind == "aziz":
        # He-4 pairs between 4.5 and 11 bohr, one pair past the cutoff
        L = 24.0
        kw.update(size=L, cutoff=0.5 * L if pbc else 9.0)
        base = np.array([[0.0, 0.0, 0.0], [6.1, 0.7, -0.4], [-4.0, 3.3, 1.9]])[:n]
        if pbc:
            base[n - 1] += np.array([L, -L, 0.0])   # its images decide the distances
    else:
        kw.update(size=30.0, cutoff=-1.0)
        base = np.array([[0.3, -0.2, 0.5], [2.9, 1.1, -0.8], [-1.7, 2.4, 1.3]])[:n]
    x = np.stack([base + rng.normal(0.0, 0.2, size=base.shape) for _ in range(P)])
    return kw, x


def pair_terms(kind, d, kw):
    """V and gradV of one pair displacement d (closed forms; Aziz through the
    oracle's pinned potential, orc_aziz_V / orc_aziz_grad)."""
    r = float(np.linalg.norm(d))
    if kind == "harmonic":
        kp = kw["mass"] * kw["int_omega"] ** 2     # the particle mass (quirk Q4)
        return 0.5 * kp * r * r, kp * d
    if kind == "dipole":
        s = kw["int_strength"]
        return s / r ** 3, -3.0 * s * d / r ** 5
    g = np.zeros(3)
    dd = np.ascontiguousarray(d, dtype=np.float64)
    O.lib().orc_aziz_grad(3, O.ptr(dd), O.ptr(g))
    return O.lib().orc_aziz_V(3, O.ptr(dd)), g


@pytest.mark.parametrize("kind,n,pbc", [("harmonic", 2, 0), ("harmonic", 3, 0), ("dipole", 3, 0),
                                        ("aziz", 3, 0), ("aziz", 3, 1)])
def test_pair_observables_worked_by_hand(kind, n, pbc):
    kw, x = pair_case(kind, n, pbc, seed=n + 7 * pbc)
    P = x.shape[0]
    kt = kw["mass"] * kw["ext_omega"] ** 2
    L = kw["size"]
    cut = kw["cutoff"]
    if pbc:
        cut = min(cut, 0.5 * L) if cut > 0 else 0.5 * L
    ext = intp = vir = 0.0
    vir_both = 0.0
    inside = 0
    for b in range(P):
        xb = x[b]
        ext += 0.5 * kt * float(np.sum(xb * xb))
        vir += kt * float(np.sum(xb * xb))            # -x . F_ext, F_ext = -k x
        vir_both += kt * float(np.sum(xb * xb))
        for i in range(n):
            for j in range(i + 1, n):
                d = xb[i] - xb[j]
                if pbc:
                    d = d - L * np.floor(d / L + 0.5)
                if cut >= 0.0 and not np.linalg.norm(d) < cut:
                    continue
                inside += 1
                v, g = pair_terms(kind, d, kw)
                intp += v
                vir -= float(xb[i] @ (-g))             # -x_i . F_{i<-j}
                vir_both -= float(xb[i] @ (-g)) + float(xb[j] @ g)
    if kind == "aziz" and not pbc:
        assert 0 < inside < P * n * (n - 1) // 2      # some pair lies past the cutoff
    sim = O.Sim(O.make_params(**kw))
    sim.set_state(x, np.zeros_like(x))
    out = sim.observables()
    potential, ext_pot, int_pot, virial = out[1], out[2], out[3], out[4]
    assert abs(ext_pot - ext / P) <= 1e-13 * abs(ext / P)
    assert abs(int_pot - intp / P) <= 1e-12 * abs(intp / P)
    assert abs(potential - (ext + intp) / P) <= 1e-12 * abs((ext + intp) / P)
    assert abs(virial - 0.5 / P * vir) <= 1e-12 * abs(0.5 / P * vir)
    # the two-ended virial differs: the test distinguishes them
    assert abs(0.5 / P * vir_both - 0.5 / P * vir) > 1e-6 * abs(0.5 / P * vir)
