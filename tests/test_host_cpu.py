"""Host C++ side of the GPU path (pimd-b_amd/host, include/pimdb_host.h) on
the CPU: unit conversion and RanMars against the reference's own compiled
units.cpp and random_mars.cpp (oracle/_ref), the configuration reader against
the reference's regression inputs, the initial state against the reference's
printed frame 0, and the noise streams against the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from helpers import ANGSTROM, GOLDEN, VEL_A_PS, case_params, load_fixture, rel_err

H = pytest.importorskip("pimdb_amd.host")

REF_LIB = os.path.join(os.path.dirname(O.__file__), "_ref", "libpimdb_ref.so")
CASES = ["bosonic_quadratic_harmonic_dynamics", "bosonic_quadratic_harmonic_nmthermostat_dynamics",
         "dist_harmonic_dynamics", "dist_harmonic_nm_propagation_dynamics"]


def ini(case):
    return os.path.join(GOLDEN, "cases", case + ".ini")


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    L = C.CDLL(REF_LIB)
    L.ref_ranmars_new.restype = C.c_void_p
    L.ref_ranmars_new.argtypes = [C.c_int]
    L.ref_ranmars_free.argtypes = [C.c_void_p]
    for n in ("ref_ranmars_uniform", "ref_ranmars_gaussian"):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_void_p]
    L.ref_units_to_internal.restype = C.c_double
    L.ref_units_to_internal.argtypes = [C.c_char_p, C.c_char_p, C.c_double]
    return L


UNITS = [("energy", "electronvolt"), ("energy", "millielectronvolt"), ("energy", "kelvin"),
         ("energy", "j/mol"), ("energy", "kilocal/mol"), ("temperature", "kelvin"), ("temperature", "millikelvin"),
         ("time", "femtosecond"), ("time", "picosecond"), ("time", "atomic_unit"), ("length", "angstrom"),
         ("length", "nanometer"), ("length", "picometer"), ("length", "degree"), ("mass", "dalton"),
         ("mass", "amu"), ("mass", "electronmass"), ("velocity", "angstrom/ps"), ("velocity", "m/s"),
         ("force", "ev/ang"), ("force", "newton"), ("frequency", "hertz*rad"), ("frequency", "inversecm"),
         ("pressure", "bar"), ("pressure", "megapascal"), ("density", "g/cm3"), ("volume", "angstrom3"),
         ("hessian", "ev/ang^2"), ("energy", ""), ("length", "automatic"), ("undefined", "atomic_unit"),
         ("energy", "millimillielectronvolt"), ("number", "anything")]


@pytest.mark.parametrize("family,unit", UNITS)
def test_units_match_reference_build(family, unit):
    L = _ref()
    for v in (1.0, 3.7, -2.25e-3):
        a = H.to_internal(family, unit, v)
        b = L.ref_units_to_internal(family.encode(), unit.encode(), v)
        assert a == b, (family, unit, v, a, b)


def test_units_errors():
    with pytest.raises(H.PimdbError, match="undefined unit for kind energy"):
        H.to_internal("energy", "furlong", 1.0)
    with pytest.raises(H.PimdbError, match="undefined units kind"):
        H.to_internal("charm", "kelvin", 1.0)


@pytest.mark.parametrize("seed", [1, 1234, 90846, 54217138, 899999999])
def test_ranmars_matches_reference_build(seed):
    L = _ref()
    h = L.ref_ranmars_new(seed)
    r = H.RanMars(seed)
    try:
        for k in range(20000):
            if k % 3:
                a, b = r.gaussian(), L.ref_ranmars_gaussian(h)
            else:
                a, b = r.uniform(), L.ref_ranmars_uniform(h)
            assert a == b, (k, a, b)
    finally:
        L.ref_ranmars_free(h)


def test_ranmars_rejects_bad_seeds():
    for s in (0, -5, 900000001):
        with pytest.raises(ValueError):
            H.RanMars(s)


@pytest.mark.parametrize("case", CASES + ["bosonic_quadratic_harmonic", "dist_harmonic"])
def test_config_matches_case_inputs(case):
    p, steps, sfreq = H.load_config(ini(case))
    kw, sf = case_params(case)
    assert (p.natoms, p.nbeads, p.ndim) == (kw["natoms"], kw["nbeads"], kw["ndim"])
    for k in ("mass", "temperature", "dt", "gamma", "size", "ext_omega"):
        assert abs(getattr(p, k) - kw[k]) <= 1e-15 * abs(kw[k]), k
    assert (p.bosonic, p.pbc, p.fixcom, p.nmthermostat, p.seed) == (
        kw["bosonic"], kw["pbc"], kw["fixcom"], kw["nmthermostat"], kw["seed"])
    assert sfreq == sf and steps > 0


@pytest.mark.parametrize("case", CASES)
def test_initial_state_matches_reference_frame0(case):
    """Positions and velocities the reference printed at step 0 (13 significant
    digits) from mt19937(seed + bead) and libstdc++'s distributions."""
    x, p = H.initial_state(ini(case))
    kw, _ = case_params(case)
    fx = load_fixture(case)
    assert rel_err(x / ANGSTROM, fx["positions"][0]) < 1e-11
    assert rel_err(p / kw["mass"] / VEL_A_PS, fx["velocities"][0]) < 1e-11


@pytest.mark.parametrize("case", ["bosonic_quadratic_harmonic_dynamics",
                                  "bosonic_quadratic_harmonic_nmthermostat_dynamics"])
def test_noise_streams_match_oracle(case):
    kw, _ = case_params(case)
    sim = O.Sim(O.make_params(**kw))
    ref = np.stack([sim.draw_noise() for _ in range(5)])
    got = H.draw_noise(ini(case), 5)
    assert np.array_equal(got.reshape(ref.shape), ref)


BASE = """[simulation]
dt = 1.0 femtosecond
steps = 100
sfreq = 10
nbeads = 4
seed = 77
thermostat = langevin
[system]
temperature = 2.0 kelvin
natoms = 3
size = 10.0 angstrom
mass = 4.0026 dalton
"""


def write_cfg(tmp_path, text, name="config.ini"):
    path = tmp_path / name
    path.write_text(text)
    return str(path)


@pytest.mark.parametrize("patch,msg", [
    (("thermostat = langevin", ""), "Thermostat must be specified!"),
    (("temperature = 2.0 kelvin", "temperature = -1.0 kelvin"), "is unphysical"),
    (("natoms = 3", "natoms = 0"), "smaller than one"),
    (("nbeads = 4", "nbeads = 0"), "less than one"),
    (("thermostat = langevin", "thermostat = langevin\nnmthermostat = true\nbosonic = true\npropagator = normal_modes"),
     "Normal modes propogation is currently not available for bosons!"),
    (("thermostat = langevin", "thermostat = none\nnmthermostat = true"), "nmthermostat cannot be used in nve ensemble!"),
    (("thermostat = langevin", "thermostat = langevin\nnchains = 3"), "nchains can only be used with Nose-Hoover"),
    (("mass = 4.0026 dalton", "mass = 4.0026 furlongs"), "undefined unit"),
    (("seed = 77", "seed = 77\ninitial_position = lattice"), "coordinate initialization method (lattice)"),
])
def test_config_validation_messages(tmp_path, patch, msg):
    text = BASE.replace(*patch)
    with pytest.raises(H.PimdbError) as ei:
        H.load_config(write_cfg(tmp_path, text))
    assert msg in str(ei.value)


def test_config_reader_syntax(tmp_path):
    """Case-insensitive keys, inline ';' comments, '#' comment lines, a ':'
    separator, a continuation line and defaults (params.cpp:14-35)."""
    text = BASE.replace("nbeads = 4", "NBeads : 6   ; six beads\n# comment").replace(
        "[system]", "fixcom = off\npbc = YES\n[System]")
    p, steps, sfreq = H.load_config(write_cfg(tmp_path, text))
    assert p.nbeads == 6 and p.fixcom == 0 and p.pbc == 1
    assert abs(p.gamma - 1.0 / (100.0 * 41.341373)) < 1e-18
    assert steps == 100 and sfreq == 10


def test_initial_state_grid_and_xyz(tmp_path):
    """grid: particles at the centres of a cubic grid (simulation.cpp:133-193);
    xyz: one frame in angstrom replicated to all beads."""
    text = BASE.replace("seed = 77", "seed = 77\ninitial_position = grid").replace("natoms = 3", "natoms = 8")
    x, _ = H.initial_state(write_cfg(tmp_path, text))
    L = 10.0 * ANGSTROM
    centres = np.array([[(i + 0.5) * L / 2 - L / 2, (j + 0.5) * L / 2 - L / 2, (k + 0.5) * L / 2 - L / 2]
                        for i in range(2) for j in range(2) for k in range(2)])
    for b in range(4):
        assert np.allclose(x[b], centres, rtol=0, atol=1e-12)
    xyz = tmp_path / "init.xyz"
    xyz.write_text("3\ncomment\nHe 0.1 0.2 0.3\nHe -1.0 2.0 0.5\nHe 3.0 -0.25 1.5\n")
    text = BASE.replace("seed = 77", "seed = 77\ninitial_position = xyz(%s)" % xyz)
    x, p = H.initial_state(write_cfg(tmp_path, text, "c2.ini"))
    want = np.array([[0.1, 0.2, 0.3], [-1.0, 2.0, 0.5], [3.0, -0.25, 1.5]]) * ANGSTROM
    for b in range(4):
        assert np.array_equal(x[b], want)
    assert abs(p.sum(axis=(0, 1))).max() < 1e-12   # Maxwell-Boltzmann momenta have their COM removed
