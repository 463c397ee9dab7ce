"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP path, never as
the thing measured or shipped.  The restatement itself is pimdb_oracle.c,
which cites the reference file:line of every function.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DP = C.POINTER(C.c_double)

POT = {"free": 0, "harmonic": 1, "dipole": 2, "aziz": 3, "double_well": 4, "cosine": 5}
PROP = {"cartesian": 0, "normal_modes": 1}
THERMO = {"none": 0, "langevin": 1, "nose_hoover": 2, "nose_hoover_np": 3, "nose_hoover_np_dim": 4}
EXCHANGE = {"quadratic": 0, "factorial": 1}


class Params(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("nbeads", C.c_int), ("ndim", C.c_int),
        ("mass", C.c_double), ("temperature", C.c_double),
        ("dt", C.c_double), ("gamma", C.c_double),
        ("bosonic", C.c_int), ("pbc", C.c_int), ("size", C.c_double),
        ("ext_kind", C.c_int), ("ext_omega", C.c_double),
        ("int_kind", C.c_int), ("int_omega", C.c_double),
        ("int_strength", C.c_double), ("cutoff", C.c_double),
        ("propagator", C.c_int), ("thermostat", C.c_int),
        ("nmthermostat", C.c_int), ("fixcom", C.c_int),
        ("seed", C.c_uint),
        ("ext_strength", C.c_double), ("ext_location", C.c_double),
        ("ext_amplitude", C.c_double), ("ext_phase", C.c_double),
        ("exchange_algo", C.c_int), ("nchains", C.c_int),
    ]


def make_params(**kw):
    """Builds Params from keyword values in atomic units; potential, propagator
    and thermostat kinds may be given by name."""
    p = Params()
    p.ndim = 3
    p.thermostat = 1
    p.cutoff = -1.0
    p.nchains = 4
    for k, v in kw.items():
        if k in ("ext_kind", "int_kind") and isinstance(v, str):
            v = POT[v]
        if k == "propagator" and isinstance(v, str):
            v = PROP[v]
        if k == "thermostat" and isinstance(v, str):
            v = THERMO[v]
        if k == "exchange_algo" and isinstance(v, str):
            v = EXCHANGE[v]
        setattr(p, k, v)
    return p


def _load():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
    lib = C.CDLL(path)
    i, d, v = C.c_int, C.c_double, None
    sigs = {
        "orc_min_image": (d, [d, d]),
        "orc_cycle_energies": (v, [i, i, _DP, _DP, d, i, d, _DP]),
        "orc_v_forward": (i, [i, d, _DP, _DP]),
        "orc_v_backward": (i, [i, d, _DP, _DP, _DP]),
        "orc_connection_probabilities": (v, [i, d, _DP, _DP, _DP, _DP]),
        "orc_spring_force_first_bead": (v, [i, i, _DP, _DP, _DP, _DP, d, i, d, _DP]),
        "orc_spring_force_last_bead": (v, [i, i, _DP, _DP, _DP, _DP, d, i, d, _DP]),
        "orc_prim_estimator": (d, [i, i, d, _DP, _DP]),
        "orc_distinct_probability": (d, [i, d, _DP, _DP]),
        "orc_longest_probability": (d, [i, d, _DP, _DP]),
        "orc_exchange": (i, [i, i, d, d, i, d, _DP, _DP, _DP, _DP,
                             _DP, _DP, _DP, _DP, _DP, _DP]),
        "orc_vb_enumerate": (d, [i, d, _DP]),
        "orc_exchange_quad": (i, [i, i, d, d, i, d, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
        "orc_factorial_e_shift": (d, [i, i, d, i, d, _DP, _DP]),
        "orc_factorial_vb": (d, [i, i, d, d, i, d, _DP, _DP]),
        "orc_factorial_force_last": (v, [i, i, d, d, i, d, _DP, _DP, _DP, d, _DP]),
        "orc_factorial_force_first": (v, [i, i, d, d, i, d, _DP, _DP, _DP, d, _DP]),
        "orc_factorial_prim_estimator": (d, [i, i, i, d, d, i, d, _DP, _DP, d]),
        "orc_harmonic_grad": (v, [i, d, d, _DP, _DP]),
        "orc_harmonic_V": (d, [i, d, d, _DP]),
        "orc_dipole_grad": (v, [i, d, _DP, _DP]),
        "orc_dipole_V": (d, [i, d, _DP]),
        "orc_aziz_grad": (v, [i, _DP, _DP]),
        "orc_aziz_V": (d, [i, _DP]),
        "orc_double_well_grad": (v, [i, d, d, d, _DP, _DP]),
        "orc_double_well_V": (d, [i, i, d, d, d, _DP]),
        "orc_cosine_grad": (v, [i, d, d, d, _DP, _DP]),
        "orc_cosine_V": (d, [i, i, d, d, d, _DP]),
        "orc_effective_cutoff": (d, [C.POINTER(Params)]),
        "orc_physical_forces": (v, [C.POINTER(Params), _DP, _DP]),
        "orc_update_forces": (i, [C.POINTER(Params), _DP, _DP]),
        "orc_nm_matrices": (v, [i, _DP, _DP]),
        "orc_sim_create": (C.c_void_p, [C.POINTER(Params)]),
        "orc_sim_destroy": (v, [C.c_void_p]),
        "orc_sim_init_random": (v, [C.c_void_p]),
        "orc_sim_set_state": (v, [C.c_void_p, _DP, _DP]),
        "orc_sim_get_state": (v, [C.c_void_p, _DP, _DP, _DP]),
        "orc_sim_draw_noise": (v, [C.c_void_p, _DP]),
        "orc_sim_step": (i, [C.c_void_p, _DP]),
        "orc_sim_forces": (i, [C.c_void_p]),
        "orc_sim_zero_momentum": (v, [C.c_void_p]),
        "orc_sim_observables": (i, [C.c_void_p, _DP]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(RuntimeError):
    pass


def derived(p):
    """beta, thermo_beta, omega_p, spring constant, exchange beta
    (src/simulation.cpp:37-52; bosonic_exchange_base.cpp:16-18)."""
    beta = 1.0 / (1.0 * p.temperature)
    omega_p = p.nbeads / (beta * 1.0)
    k = p.mass * omega_p * omega_p
    return dict(beta=beta, thermo_beta=beta / p.nbeads, omega_p=omega_p, k=k,
                beta_t=beta / p.nbeads)


def exchange(x1, x2, xPm1, xP, beta_t, k, pbc=False, L=0.0):
    """Full exchange on N x D bead arrays: returns dict(E, V, Vb, prob, f_first, f_last)."""
    x1, x2, xPm1, xP = (f64(a) for a in (x1, x2, xPm1, xP))
    N, D = x1.shape
    out = dict(E=np.zeros(N * (N + 1) // 2), V=np.zeros(N + 1), Vb=np.zeros(N + 1),
               prob=np.zeros((N, N)), f_first=np.zeros((N, D)), f_last=np.zeros((N, D)))
    rc = lib().orc_exchange(N, D, beta_t, k, int(pbc), L, ptr(x1), ptr(x2), ptr(xPm1), ptr(xP),
                            *(ptr(out[n]) for n in ("E", "V", "Vb", "prob", "f_first", "f_last")))
    if rc != 0:
        raise OracleError("exchange returned %d" % rc)
    return out


def exchange_quad(x1, x2, xPm1, xP, beta_t, k, pbc=False, L=0.0):
    """orc_exchange in quadruple precision (the exact values for the double
    inputs): dict(V, Vb, prob, f_first, f_last)."""
    x1, x2, xPm1, xP = (f64(a) for a in (x1, x2, xPm1, xP))
    N, D = x1.shape
    out = dict(V=np.zeros(N + 1), Vb=np.zeros(N + 1), prob=np.zeros((N, N)), f_first=np.zeros((N, D)),
               f_last=np.zeros((N, D)))
    rc = lib().orc_exchange_quad(N, D, beta_t, k, int(pbc), L, ptr(x1), ptr(x2), ptr(xPm1), ptr(xP),
                                 *(ptr(out[n]) for n in ("V", "Vb", "prob", "f_first", "f_last")))
    if rc != 0:
        raise OracleError("exchange_quad returned %d" % rc)
    return out


def cycle_energies(x1, xP, k, pbc=False, L=0.0):
    x1, xP = f64(x1), f64(xP)
    N, D = x1.shape
    E = np.zeros(N * (N + 1) // 2)
    lib().orc_cycle_energies(N, D, ptr(x1), ptr(xP), k, int(pbc), L, ptr(E))
    return E


def vb_enumerate(E, N, beta_t):
    return lib().orc_vb_enumerate(N, beta_t, ptr(f64(E)))


def update_forces(p, x):
    x = f64(x)
    f = np.zeros_like(x)
    rc = lib().orc_update_forces(C.byref(p), ptr(x), ptr(f))
    if rc != 0:
        raise OracleError("update_forces returned %d" % rc)
    return f


def physical_forces(p, coord):
    coord = f64(coord)
    f = np.zeros_like(coord)
    lib().orc_physical_forces(C.byref(p), ptr(coord), ptr(f))
    return f


def nm_matrices(P):
    Cm = np.zeros((P, P))
    CT = np.zeros((P, P))
    lib().orc_nm_matrices(P, ptr(Cm), ptr(CT))
    return Cm, CT


class Sim:
    """Whole-ring oracle simulation ([bead][particle][axis] state)."""

    def __init__(self, p):
        self.p = p
        self.h = lib().orc_sim_create(C.byref(p))
        if not self.h:
            raise OracleError("bad parameters")
        self.shape = (p.nbeads, p.natoms, p.ndim)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def init_random(self):
        lib().orc_sim_init_random(self.h)

    def set_state(self, x=None, p=None):
        lib().orc_sim_set_state(self.h, ptr(None if x is None else f64(x)),
                                ptr(None if p is None else f64(p)))

    def state(self):
        x, p, f = (np.zeros(self.shape) for _ in range(3))
        lib().orc_sim_get_state(self.h, ptr(x), ptr(p), ptr(f))
        return x, p, f

    def draw_noise(self):
        n = np.zeros((2,) + self.shape)
        lib().orc_sim_draw_noise(self.h, ptr(n))
        return n

    def step(self, noise=None):
        rc = lib().orc_sim_step(self.h, ptr(None if noise is None else f64(noise)))
        if rc != 0:
            raise OracleError("step returned %d" % rc)

    def forces(self):
        rc = lib().orc_sim_forces(self.h)
        if rc != 0:
            raise OracleError("forces returned %d" % rc)

    def observables(self):
        out = np.zeros(11)
        rc = lib().orc_sim_observables(self.h, ptr(out))
        if rc != 0:
            raise OracleError("observables returned %d" % rc)
        return out


def factorial_exchange(N, D, beta_t, k, pbc, L, x_first, x_last, x_second, x_prev_last):
    """The factorial exchange (factorial_bosonic_exchange.cpp) on AoS [N][D]
    beads: e_shift, V_eff, forces on beads 1 and P, kinetic estimator term
    (before the /P of the IPI convention is applied by the caller: returned
    with P = 1)."""
    x1, xP, x2, xPm1 = (f64(a) for a in (x_first, x_last, x_second, x_prev_last))
    L_ = lib()
    es = L_.orc_factorial_e_shift(N, D, k, pbc, L, ptr(x1), ptr(xP))
    vb = L_.orc_factorial_vb(N, D, beta_t, k, pbc, L, ptr(x1), ptr(xP))
    f1 = np.zeros((N, D))
    fP = np.zeros((N, D))
    L_.orc_factorial_force_first(N, D, beta_t, k, pbc, L, ptr(x1), ptr(xP), ptr(x2), es, ptr(f1))
    L_.orc_factorial_force_last(N, D, beta_t, k, pbc, L, ptr(x1), ptr(xP), ptr(xPm1), es, ptr(fP))
    prim = L_.orc_factorial_prim_estimator(N, 1, D, beta_t, k, pbc, L, ptr(x1), ptr(xP), es)
    return dict(e_shift=es, vb=vb, f_first=f1, f_last=fP, prim=prim)
