"""The oracle against what the mathematics fixes: brute-force enumeration of
V_B, closed forms, invariants of the connection probabilities, finite
differences of V_B and of the pair potentials, image enumeration for the
minimum image and the cutoff, and the normal-mode transform."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import ANGSTROM, KELVIN


def ring(N, D, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [rng.normal(0, scale, size=(N, D)) for _ in range(4)]


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8, 12])
@pytest.mark.parametrize("pbc", [False, True])
def test_vb_matches_composition_enumeration(N, pbc):
    x1, x2, xPm1, xP = ring(N, 3, 100 + N, scale=1.5)
    bt, k = 2.0, 0.7
    out = O.exchange(x1, x2, xPm1, xP, bt, k, pbc, 4.0)
    vb = O.vb_enumerate(out["E"], N, bt)
    assert abs(out["V"][N] - vb) <= 1e-12 * max(1.0, abs(vb))


def test_cycle_energy_closed_form():
    N, D = 9, 2
    x1, _, _, xP = ring(N, D, 5)
    k = 1.3
    E = O.cycle_energies(x1, xP, k)
    for v in range(N):
        for u in range(v + 1):
            s = sum(np.sum((x1[j + 1] - xP[j]) ** 2) for j in range(u, v)) + np.sum((x1[u] - xP[v]) ** 2)
            m, kk = v + 1, v - u + 1
            assert abs(E[m * (m + 1) // 2 - kk] - 0.5 * k * s) < 1e-12 * max(1, 0.5 * k * s)


def test_closed_forms_N1_N2():
    k, bt = 0.9, 1.7
    x1, x2, xPm1, xP = ring(1, 3, 1)
    out = O.exchange(x1, x2, xPm1, xP, bt, k)
    assert abs(out["V"][1] - 0.5 * k * np.sum((x1 - xP) ** 2)) < 1e-14
    x1, x2, xPm1, xP = ring(2, 3, 2)
    out = O.exchange(x1, x2, xPm1, xP, bt, k)
    E11 = 0.5 * k * np.sum((x1[0] - xP[0]) ** 2)
    E22 = 0.5 * k * np.sum((x1[1] - xP[1]) ** 2)
    E12 = 0.5 * k * (np.sum((x1[1] - xP[0]) ** 2) + np.sum((x1[0] - xP[1]) ** 2))
    vb = -1.0 / bt * math.log(0.5 * (math.exp(-bt * (E11 + E22)) + math.exp(-bt * E12)))
    assert abs(out["V"][2] - vb) < 1e-13


@pytest.mark.parametrize("N", [2, 5, 17])
def test_coincident_beads(N):
    """All beads of all particles at one point: every cycle energy is zero."""
    rng = np.random.default_rng(N)
    x = np.repeat(rng.normal(0, 1, size=(1, 3)), N, axis=0)
    out = O.exchange(x, x, x, x, 3.0, 1.1)
    assert np.max(np.abs(out["V"])) < 1e-13
    P = out["prob"]
    for l in range(N - 1):
        assert np.allclose(P[l, :l + 1], 1.0 / ((l + 1) * (l + 2)), rtol=1e-13, atol=0)
        assert abs(P[l, l + 1] - (l + 1) / (l + 2)) < 1e-13
    assert np.allclose(P[N - 1], 1.0 / N, rtol=1e-13, atol=0)


def test_distinguishable_limit():
    N, bt, k = 6, 1.0, 1.0
    rng = np.random.default_rng(3)
    centres = np.arange(N)[:, None] * np.array([[40.0, 0.0, 0.0]])
    x1 = centres + rng.normal(0, 0.1, size=(N, 3))
    xP = centres + rng.normal(0, 0.1, size=(N, 3))
    out = O.exchange(x1, x1, xP, xP, bt, k)
    expect = sum(0.5 * k * np.sum((x1[i] - xP[i]) ** 2) for i in range(N)) + math.lgamma(N + 1) / bt
    assert abs(out["V"][N] - expect) < 1e-10
    assert np.allclose(out["prob"], np.eye(N), atol=1e-12)


@pytest.mark.parametrize("pbc", [False, True])
def test_probability_sums(pbc):
    N = 11
    x1, x2, xPm1, xP = ring(N, 3, 9, scale=0.8)
    out = O.exchange(x1, x2, xPm1, xP, 1.3, 0.8, pbc, 3.0)
    P = out["prob"]
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-12)
    assert np.allclose(P.sum(axis=0), 1.0, atol=1e-12)


@pytest.mark.parametrize("pbc", [False, True])
def test_exchange_forces_are_minus_gradient(pbc):
    """F1, FP = -grad(V_B + k/2|x2-x1|^2 + k/2|x_{P-1}-x_P|^2) by central FD
    (independent of the probability formulas)."""
    N, D, bt, k, L = 6, 3, 1.1, 0.9, 5.0
    x1, x2, xPm1, xP = ring(N, D, 21, scale=1.2)

    def energy(a1, aP):
        V = O.exchange(a1, x2, xPm1, aP, bt, k, pbc, L)["V"][N]
        d1 = x2 - a1
        dP = xPm1 - aP
        if pbc:
            d1 = d1 - L * np.floor(d1 / L + 0.5)
            dP = dP - L * np.floor(dP / L + 0.5)
        return V + 0.5 * k * np.sum(d1 ** 2) + 0.5 * k * np.sum(dP ** 2)

    out = O.exchange(x1, x2, xPm1, xP, bt, k, pbc, L)
    h = 1e-6
    for which, f in ((0, out["f_first"]), (1, out["f_last"])):
        for i in range(N):
            for a in range(D):
                xa, xb = x1.copy(), xP.copy()
                tgt = xa if which == 0 else xb
                tgt[i, a] += h
                ep = energy(xa, xb)
                tgt[i, a] -= 2 * h
                em = energy(xa, xb)
                fd = -(ep - em) / (2 * h)
                assert abs(fd - f[i, a]) < 2e-7 * max(1.0, abs(f[i, a])), (which, i, a, fd, f[i, a])


def _grad(kind, d, **kw):
    L = O.lib()
    g = np.zeros(len(d))
    d = np.ascontiguousarray(d, dtype=float)
    if kind == "aziz":
        L.orc_aziz_grad(len(d), O.ptr(d), O.ptr(g))
    elif kind == "dipole":
        L.orc_dipole_grad(len(d), kw["s"], O.ptr(d), O.ptr(g))
    else:
        L.orc_harmonic_grad(len(d), kw["m"], kw["w"], O.ptr(d), O.ptr(g))
    return g


def _pot(kind, d, **kw):
    L = O.lib()
    d = np.ascontiguousarray(d, dtype=float)
    if kind == "aziz":
        return L.orc_aziz_V(len(d), O.ptr(d))
    if kind == "dipole":
        return L.orc_dipole_V(len(d), kw["s"], O.ptr(d))
    return L.orc_harmonic_V(len(d), kw["m"], kw["w"], O.ptr(d))


@pytest.mark.parametrize("r", [0.04, 2.5, 4.0, 5.60738, 6.9, 9.0, 15.0])
def test_aziz_gradient_matches_fd(r):
    rng = np.random.default_rng(int(r * 100))
    u = rng.normal(size=3)
    d = r * u / np.linalg.norm(u)
    g = _grad("aziz", d)
    h = 1e-6 * max(r, 1.0)
    for a in range(3):
        dp, dm = d.copy(), d.copy()
        dp[a] += h
        dm[a] -= h
        fd = (_pot("aziz", dp) - _pot("aziz", dm)) / (2 * h)
        assert abs(fd - g[a]) <= 1e-6 * np.max(np.abs(g)) + 1e-14


def test_aziz_well_depth():
    """HFDHE2: V(r_m) ~ -epsilon (10.8 K at 2.9673 angstrom)."""
    eps = 3.42016e-5
    v = _pot("aziz", np.array([5.60738, 0.0, 0.0]))
    assert abs(v / eps + 1.0) < 1e-3
    assert abs(v / KELVIN + 10.8) < 0.02
    assert abs(5.60738 / ANGSTROM - 2.9673) < 1e-4


def test_harmonic_and_dipole_closed_forms():
    d = np.array([0.3, -1.2, 2.0])
    r = np.linalg.norm(d)
    assert np.allclose(_grad("harmonic", d, m=3.0, w=0.5), 3.0 * 0.25 * d, rtol=1e-15)
    assert abs(_pot("harmonic", d, m=3.0, w=0.5) - 0.5 * 0.75 * r * r) < 1e-14
    assert np.allclose(_grad("dipole", d, s=2.5), -3 * 2.5 / r ** 5 * d, rtol=1e-14)
    assert abs(_pot("dipole", d, s=2.5) - 2.5 / r ** 3) < 1e-14


def test_min_image_brute_force():
    rng = np.random.default_rng(0)
    L = 7.3
    for dx in rng.uniform(-40, 40, size=2000):
        m = O.lib().orc_min_image(dx, L)
        best = min((abs(dx - n * L) for n in range(-10, 11)))
        assert abs(abs(m) - best) < 1e-12
        assert -0.5 * L <= m <= 0.5 * L


@pytest.mark.parametrize("kind", ["aziz", "dipole", "harmonic"])
def test_pair_forces_image_enumeration(kind):
    """Physical forces with PBC and cutoff L/2 against an explicit search of
    the nearest image of every pair and the cutoff rule |r| < c."""
    N, D, L = 7, 3, 14.0
    rng = np.random.default_rng(11)
    x = rng.uniform(-L, L, size=(N, D))
    kw = dict(natoms=N, nbeads=1, ndim=D, mass=2.0, temperature=1.0, dt=1.0, gamma=0.0,
              bosonic=0, pbc=1, size=L, ext_kind="harmonic", ext_omega=0.03,
              int_kind=kind, int_omega=0.02, int_strength=0.7, cutoff=100.0)  # clamped to L/2
    p = O.make_params(**kw)
    f = O.physical_forces(p, x)
    ref = -2.0 * 0.03 ** 2 * x
    shifts = [np.array(s) for s in np.ndindex(5, 5, 5)]
    for i in range(N):
        for j in range(N):
            if i == j:
                continue
            cands = [x[i] - x[j] - L * (s - 2) for s in shifts]
            d = min(cands, key=lambda c: np.linalg.norm(c))
            if np.linalg.norm(d) < 0.5 * L:
                ref[i] -= _grad(kind, d, s=0.7, m=2.0, w=0.02)
    assert np.max(np.abs(f - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("P", [1, 2, 3, 8, 9, 64])
def test_nm_matrices_orthogonal(P):
    Cm, CT = O.nm_matrices(P)
    assert np.allclose(Cm @ Cm.T, np.eye(P), atol=1e-12)
    assert np.allclose(CT, Cm.T, atol=1e-12)


def test_nm_propagator_free_ring_is_exact():
    """Free distinguishable ring, no thermostat: every normal mode rotates
    exactly at w_k = 2 w_P sin(k pi/P) (normal_modes_propagator.cpp:13-39)."""
    N, P, D = 3, 8, 3
    kw = dict(natoms=N, nbeads=P, ndim=D, mass=5.0, temperature=0.01, dt=2.0, gamma=0.0,
              bosonic=0, pbc=0, size=10.0, ext_kind="free", int_kind="free",
              propagator="normal_modes", thermostat="none", fixcom=0, seed=5)
    p = O.make_params(**kw)
    rng = np.random.default_rng(1)
    x0 = rng.normal(0, 1, size=(P, N, D))
    p0 = rng.normal(0, 1, size=(P, N, D))
    sim = O.Sim(p)
    sim.set_state(x0, p0)
    nsteps = 50
    for _ in range(nsteps):
        sim.step()
    x, pm, _ = sim.state()
    Cm, _ = O.nm_matrices(P)
    q0 = np.einsum("kj,jnd->knd", Cm, x0)
    v0 = np.einsum("kj,jnd->knd", Cm, p0)
    beta = 1.0 / 0.01
    wp = P / beta
    t = nsteps * 2.0
    qt = np.empty_like(q0)
    for k in range(P):
        w = 2 * wp * math.sin(k * math.pi / P)
        if w == 0:
            qt[k] = q0[k] + t / 5.0 * v0[k]
        else:
            qt[k] = math.cos(w * t) * q0[k] + math.sin(w * t) / (5.0 * w) * v0[k]
    xt = np.einsum("kj,knd->jnd", Cm, qt)
    assert np.max(np.abs(x - xt)) < 1e-10


# ---- external double well and cosine (double_well.cpp, cosine.cpp) ----

def _dw(x, m=2.0, s=0.7, loc=1.3):
    L = O.lib()
    g = np.zeros(len(x))
    L.orc_double_well_grad(len(x), m, s, loc, O.ptr(x), O.ptr(g))
    return g


def test_double_well_gradient_is_isotropic_well():
    """The reference's gradV is the gradient of the isotropic well
    U(r) = m s (r^2 - loc^2)^2 (its V sums per axis instead; the force is what
    enters the dynamics).  Central differences of U at random 3-D points."""
    m, s, loc = 2.0, 0.7, 1.3
    U = lambda y: m * s * (float(y @ y) - loc * loc) ** 2
    rng = np.random.default_rng(8)
    for _ in range(5):
        x = rng.normal(0, 1.2, size=3)
        g = _dw(x, m, s, loc)
        h = 1e-6
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a] += h
            xm[a] -= h
            assert abs((U(xp) - U(xm)) / (2 * h) - g[a]) <= 1e-7 * np.max(np.abs(g)) + 1e-12


def test_double_well_energy_closed_forms():
    """Per-axis energy: V(0) = N D m s loc^4; on one axis V and its derivative
    agree with the gradient (the isotropic and per-axis forms coincide there);
    minima at +-loc along an axis with the other axes at 0 cost m s loc^4 each."""
    L = O.lib()
    m, s, loc = 2.0, 0.7, 1.3
    N, D = 3, 3
    zero = np.zeros(N * D)
    assert abs(L.orc_double_well_V(N, D, m, s, loc, O.ptr(zero)) - N * D * m * s * loc ** 4) < 1e-13
    x1 = np.array([loc, 0.0, 0.0])
    assert abs(L.orc_double_well_V(1, 3, m, s, loc, O.ptr(x1)) - 2 * m * s * loc ** 4) < 1e-13
    for a0 in (0.4, -1.7, 2.2):
        h = 1e-6
        vp = L.orc_double_well_V(1, 1, m, s, loc, O.ptr(np.array([a0 + h])))
        vm = L.orc_double_well_V(1, 1, m, s, loc, O.ptr(np.array([a0 - h])))
        g = _dw(np.array([a0]), m, s, loc)[0]
        assert abs((vp - vm) / (2 * h) - g) <= 1e-7 * abs(g) + 1e-12


def test_cosine_potential():
    """V = A sum cos(k x + phase), k = 2 pi / L: gradient is the derivative of
    V (central differences), V = A D at k x + phase = 0 on every axis, and V is
    periodic in the box."""
    L = O.lib()
    A, Lbox, ph = 0.02, 7.5, 0.3
    k = 2 * np.pi / Lbox
    rng = np.random.default_rng(4)
    x = rng.uniform(-Lbox, Lbox, size=3)
    g = np.zeros(3)
    L.orc_cosine_grad(3, A, Lbox, ph, O.ptr(x), O.ptr(g))
    h = 1e-6
    for a in range(3):
        xp, xm = x.copy(), x.copy()
        xp[a] += h
        xm[a] -= h
        fd = (L.orc_cosine_V(1, 3, A, Lbox, ph, O.ptr(xp)) - L.orc_cosine_V(1, 3, A, Lbox, ph, O.ptr(xm))) / (2 * h)
        assert abs(fd - g[a]) <= 1e-7 * np.max(np.abs(g)) + 1e-13
    x0 = np.full(3, -ph / k)
    assert abs(L.orc_cosine_V(1, 3, A, Lbox, ph, O.ptr(x0)) - 3 * A) < 1e-15
    xs = x + Lbox
    assert abs(L.orc_cosine_V(1, 3, A, Lbox, ph, O.ptr(xs)) - L.orc_cosine_V(1, 3, A, Lbox, ph, O.ptr(x))) < 1e-14


def test_physical_forces_use_external_gradient():
    """orc_physical_forces with a double-well or cosine trap and free pairs is
    -gradV_ext of every particle (simulation.cpp:430)."""
    rng = np.random.default_rng(12)
    x = rng.normal(0, 1.0, size=(6, 3))
    for ext in ("double_well", "cosine"):
        kw = dict(natoms=6, nbeads=1, ndim=3, mass=2.0, temperature=1e-5, dt=1.0, gamma=0.0, bosonic=0,
                  pbc=0, size=7.5, ext_kind=ext, int_kind="free", ext_strength=0.7, ext_location=1.3,
                  ext_amplitude=0.02, ext_phase=0.3)
        f = np.zeros((6, 3))
        O.lib().orc_physical_forces(O.make_params(**kw), O.ptr(x), O.ptr(f))
        for i in range(6):
            g = np.zeros(3)
            if ext == "double_well":
                O.lib().orc_double_well_grad(3, 2.0, 0.7, 1.3, O.ptr(x[i].copy()), O.ptr(g))
            else:
                O.lib().orc_cosine_grad(3, 0.02, 7.5, 0.3, O.ptr(x[i].copy()), O.ptr(g))
            assert np.array_equal(f[i], -g)


# ---- factorial exchange (factorial_bosonic_exchange.cpp) ----

def _perm_energies(x1, xP, k, pbc, L):
    """0.5 k sum_l |x_P[sigma(l)] - x1[l]|^2 for every permutation sigma,
    by itertools (independent of the oracle's next_permutation walk)."""
    import itertools
    N = x1.shape[0]
    out = []
    for s in itertools.permutations(range(N)):
        d = xP[list(s)] - x1
        if pbc:
            d = d - L * np.floor(d / L + 0.5)
        out.append(0.5 * k * np.sum(d * d))
    return np.array(out)


@pytest.mark.parametrize("pbc", [False, True])
def test_factorial_vb_and_shift_brute_force(pbc):
    """V_eff = -(1/beta~) ln(mean_sigma exp(-beta~ E_sigma)) and e_shift =
    min_sigma E_sigma against an itertools enumeration."""
    N, D, bt, k, L = 5, 3, 0.8, 0.7, 4.0
    x1, x2, xPm1, xP = ring(N, D, 31, scale=1.1)
    out = O.factorial_exchange(N, D, bt, k, int(pbc), L, x1, xP, x2, xPm1)
    E = _perm_energies(x1, xP, k, pbc, L)
    assert abs(out["e_shift"] - E.min()) <= 1e-14 * E.min()
    vb = E.min() - np.log(np.mean(np.exp(-bt * (E - E.min())))) / bt
    assert abs(out["vb"] - vb) <= 1e-12 * abs(vb)
    # kinetic term: -<E_sigma> over the Boltzmann weights (P = 1 here)
    w = np.exp(-bt * (E - E.min()))
    assert abs(out["prim"] + np.sum(E * w) / np.sum(w)) <= 1e-12 * np.sum(E * w) / np.sum(w)


def test_factorial_equals_quadratic_for_two_particles():
    """For N = 2 the permutation average and the Feldman-Hirshberg recursion
    are the same potential (V_B, forces and the kinetic term agree)."""
    N, D, bt, k = 2, 3, 1.3, 0.9
    x1, x2, xPm1, xP = ring(N, D, 41, scale=0.9)
    fa = O.factorial_exchange(N, D, bt, k, 0, 0.0, x1, xP, x2, xPm1)
    fh = O.exchange(x1, x2, xPm1, xP, bt, k)
    assert abs(fa["vb"] - fh["V"][N]) <= 1e-12 * abs(fh["V"][N])
    assert np.max(np.abs(fa["f_first"] - fh["f_first"])) <= 1e-12 * np.max(np.abs(fh["f_first"]))
    assert np.max(np.abs(fa["f_last"] - fh["f_last"])) <= 1e-12 * np.max(np.abs(fh["f_last"]))
    E = O.cycle_energies(x1, xP, k)
    prim_fh = O.lib().orc_prim_estimator(N, 1, bt, O.ptr(E), O.ptr(fh["V"]))
    assert abs(fa["prim"] - prim_fh) <= 1e-12 * abs(prim_fh)


@pytest.mark.parametrize("pbc", [False, True])
def test_factorial_forces_are_minus_gradient(pbc):
    """F1, FP = -grad(V_eff + interior links) by central differences."""
    N, D, bt, k, L = 4, 3, 1.1, 0.9, 5.0
    x1, x2, xPm1, xP = ring(N, D, 51, scale=1.2)

    def energy(a1, aP):
        V = O.factorial_exchange(N, D, bt, k, int(pbc), L, a1, aP, x2, xPm1)["vb"]
        d1 = x2 - a1
        dP = xPm1 - aP
        if pbc:
            d1 = d1 - L * np.floor(d1 / L + 0.5)
            dP = dP - L * np.floor(dP / L + 0.5)
        return V + 0.5 * k * np.sum(d1 ** 2) + 0.5 * k * np.sum(dP ** 2)

    out = O.factorial_exchange(N, D, bt, k, int(pbc), L, x1, xP, x2, xPm1)
    h = 1e-6
    for which, f in ((0, out["f_first"]), (1, out["f_last"])):
        for i in range(N):
            for a in range(D):
                xa, xb = x1.copy(), xP.copy()
                tgt = xa if which == 0 else xb
                tgt[i, a] += h
                ep = energy(xa, xb)
                tgt[i, a] -= 2 * h
                em = energy(xa, xb)
                fd = -(ep - em) / (2 * h)
                assert abs(fd - f[i, a]) < 2e-7 * max(1.0, abs(f[i, a])), (which, i, a, fd, f[i, a])


# ---- Nose-Hoover chains (nose_hoover.cpp) ----

def _nh_extended_energy(sim, kw):
    """Ring-polymer Hamiltonian at the bead temperature (harmonic trap,
    distinguishable springs) plus the thermostats' getAdditionToH."""
    x, p, _ = sim.state()
    P = kw["nbeads"]
    d = O.derived(O.make_params(**kw))
    m, w = kw["mass"], kw["ext_omega"]
    ke = np.sum(p * p) / (2 * m)
    spring = 0.5 * d["k"] * np.sum((x - np.roll(x, 1, axis=0)) ** 2)
    pot = 0.5 * m * w * w * np.sum(x * x)
    return ke + spring + pot + sim.observables()[10]


@pytest.mark.parametrize("kind", ["nose_hoover", "nose_hoover_np", "nose_hoover_np_dim"])
@pytest.mark.parametrize("nm", [0, 1])
def test_nose_hoover_conserves_extended_energy(kind, nm):
    """The Nose-Hoover equations conserve H + getAdditionToH; its change over
    400 steps is small and shrinks as dt^2 (Cartesian and normal-mode
    coupling; the latter has no reference fixture)."""
    drifts = []
    for dt in (2.0, 1.0):
        kw = dict(natoms=3, nbeads=6, ndim=3, mass=4.0026 * 1822.8885, temperature=8e-6, dt=dt * 41.341373,
                  gamma=0.0, bosonic=0, pbc=0, size=100.0, ext_kind="harmonic", ext_omega=1.1e-4,
                  int_kind="free", cutoff=0.0, propagator="cartesian", thermostat=kind, nmthermostat=nm,
                  fixcom=0, seed=77, nchains=4)
        sim = O.Sim(O.make_params(**kw))
        sim.init_random()
        sim.forces()
        e0 = _nh_extended_energy(sim, kw)
        steps = int(400 * 2.0 / dt)
        emax = 0.0
        for _ in range(steps):
            sim.step()
            emax = max(emax, abs(_nh_extended_energy(sim, kw) - e0))
        drifts.append(emax / abs(e0))
    assert drifts[0] < 1e-3, drifts
    assert drifts[1] < 0.4 * drifts[0], drifts


@pytest.mark.parametrize("N,pbc", [(1, 0), (2, 1), (5, 0), (8, 1)])
def test_quad_exchange_pins(N, pbc):
    """oracle.exchange_quad (the __float128 restatement used as the exact
    value at N=1024): agrees with the double oracle to its rounding, its V_B
    with the composition enumeration, and its exterior forces (with the
    interior links taken out) with central differences of its own V_B."""
    rng = np.random.default_rng(40 + N)
    L, beta, k = 6.0, 1.1, 0.7
    x1 = rng.uniform(-2, 2, size=(N, 3))
    xP = x1 + rng.normal(0, 0.6, size=(N, 3))
    q = O.exchange_quad(x1, x1, xP, xP, beta, k, pbc, L)
    o = O.exchange(x1, x1, xP, xP, beta, k, pbc, L)
    for key in ("V", "Vb", "f_first", "f_last"):
        assert np.max(np.abs(q[key] - o[key])) <= 1e-13 * max(1.0, np.max(np.abs(q[key]))), key
    assert np.max(np.abs(q["prob"] - o["prob"])) <= 1e-13
    E = O.cycle_energies(x1, xP, k, pbc, L)
    assert abs(q["V"][N] - O.vb_enumerate(E, N, beta)) <= 1e-12 * max(1.0, abs(q["V"][N]))
    h = 1e-6
    for which, xs in ((0, x1), (1, xP)):
        for i in range(N):
            for a in range(3):
                xp, xm = xs.copy(), xs.copy()
                xp[i, a] += h
                xm[i, a] -= h
                if which == 0:
                    vp = O.exchange_quad(xp, xp, xP, xP, beta, k, pbc, L)["V"][N]
                    vm = O.exchange_quad(xm, xm, xP, xP, beta, k, pbc, L)["V"][N]
                    got = q["f_first"][i, a]
                else:
                    vp = O.exchange_quad(x1, x1, xp, xp, beta, k, pbc, L)["V"][N]
                    vm = O.exchange_quad(x1, x1, xm, xm, beta, k, pbc, L)["V"][N]
                    got = q["f_last"][i, a]
                assert abs(got + (vp - vm) / (2 * h)) < 1e-7, (which, i, a)
