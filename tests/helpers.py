"""Shared test helpers: unit constants of the reference (include/units.h:18-153),
fixture loading, configuration builders and error norms."""
import configparser
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

# include/units.h (i-PI values, used verbatim)
KELVIN = 3.1668152e-06
EV = 0.036749326
ANGSTROM = 1.8897261
SECOND = 4.1341373e16
AMU = 1822.8885
VEL_A_PS = 4.5710289e-5
FORCE_EV_A = 0.019446904
FS = SECOND * 1e-15
MEV = EV * 1e-3


def load_fixture(case):
    return np.load(os.path.join(GOLDEN, case + ".npz"), allow_pickle=False)


def case_params(case):
    """Keyword parameters (atomic units) of a reference regression case .ini,
    following the defaults of src/params.cpp:14-247."""
    cp = configparser.ConfigParser(inline_comment_prefixes=(";", "#"))
    cp.read(os.path.join(GOLDEN, "cases", case + ".ini"))
    sim, sys_ = cp["simulation"], cp["system"]

    def qty(s, unit_map):
        val, unit = s.split()
        return float(val) * unit_map[unit]

    time_units = {"femtosecond": FS, "atomic_unit": 1.0}
    dt = qty(sim.get("dt", "1.0 femtosecond"), time_units)
    temp = qty(sys_.get("temperature"), {"kelvin": KELVIN, "atomic_unit": 1.0})
    mass = qty(sys_.get("mass"), {"atomic_unit": 1.0, "dalton": AMU})
    size = qty(sys_.get("size"), {"atomic_unit": 1.0, "angstrom": ANGSTROM})
    ext = cp["external_potential"]
    omega = qty(ext.get("omega", "1.0 millielectronvolt"), {"millielectronvolt": MEV})
    b = lambda s, d: s.lower() in ("true", "1", "yes", "on") if s is not None else d
    return dict(
        natoms=int(sys_.get("natoms")), nbeads=int(float(sim.get("nbeads"))), ndim=3,
        mass=mass, temperature=temp, dt=dt, gamma=1.0 / (100.0 * dt),
        bosonic=int(b(sim.get("bosonic"), False)), pbc=int(b(sim.get("pbc"), False)), size=size,
        ext_kind=ext.get("name", "free"), ext_omega=omega,
        int_kind=cp["interaction_potential"].get("name", "free"), cutoff=-1.0,
        propagator=sim.get("propagator", "cartesian"), thermostat=sim.get("thermostat"),
        nmthermostat=int(b(sim.get("nmthermostat"), False)), fixcom=int(b(sim.get("fixcom"), True)),
        seed=int(float(sim.get("seed", "1234"))),
        # the reference picks the exchange algorithm at build time
        # (FACTORIAL_BOSONIC_ALGORITHM); its harness runs the *factorial* cases
        # with that build (tests/main.py:286-290)
        exchange_algo=1 if "factorial" in case else 0,
        nchains=int(sim.get("nchains", "4")),
    ), int(sim.get("sfreq", "1000"))


def rel_err(a, b):
    """Norm-wise relative error max|a-b| / max|b| (per array)."""
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def per_bead_err(a, b):
    """max over beads of max|a-b|/max|b| -- the norm-wise force error per bead."""
    return max(rel_err(a[k], b[k]) for k in range(a.shape[0]))


def aziz_config(N, P, T_kelvin, density_A3=0.02186, seed=12345, ndim=3, replicate=True):
    """He-4 Aziz configuration: PBC cube at the given number density, cutoff
    L/2, seeded uniform positions (replicated across beads if asked)."""
    L = (N / density_A3) ** (1.0 / ndim) * ANGSTROM
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-0.5 * L, 0.5 * L, size=(N, ndim))
    x = np.repeat(x0[None], P, axis=0) if replicate else rng.uniform(-0.5 * L, 0.5 * L, size=(P, N, ndim))
    kw = dict(natoms=N, nbeads=P, ndim=ndim, mass=4.0026 * AMU, temperature=T_kelvin * KELVIN,
              dt=0.5 * FS, gamma=1.0 / (100 * 0.5 * FS), bosonic=1, pbc=1, size=L,
              ext_kind="free", int_kind="aziz", cutoff=0.5 * L, propagator="cartesian",
              thermostat="langevin", nmthermostat=1, fixcom=1, seed=seed % 900000000 + 1)
    return kw, x
