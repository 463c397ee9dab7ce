"""Python bindings (ctypes) to libpimdb_gpu.so, the MI355X implementation of
the PIMD-B++ per-step hot path.  The C ABI is include/pimdb_gpu.h.

There is no CPU fallback: loading fails with an error when the HIP library is
not built, and every call that cannot reach a GPU raises PimdbError.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libpimdb_gpu.so")
_DP = C.POINTER(C.c_double)

OK, E_NONFINITE, E_ARG, E_HIP, E_STATE = 0, 1, 2, 3, 4
POT = {"free": 0, "harmonic": 1, "dipole": 2, "aziz": 3, "double_well": 4, "cosine": 5}
PROP = {"cartesian": 0, "normal_modes": 1}
THERMO = {"none": 0, "langevin": 1, "nose_hoover": 2, "nose_hoover_np": 3, "nose_hoover_np_dim": 4}
EXCHANGE = {"quadratic": 0, "factorial": 1}


class PimdbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pimdb error %d: %s" % (code, msg))
        self.code = code


class NonFiniteError(PimdbError, OverflowError):
    """Mirrors the reference's std::overflow_error from the exchange."""


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
    p = Params()
    p.ndim = 3
    p.thermostat = 1
    p.cutoff = -1.0
    p.nchains = 4
    for k, v in kw.items():
        if k in ("ext_kind", "int_kind") and isinstance(v, str):
            v = POT[v]
        elif k == "propagator" and isinstance(v, str):
            v = PROP[v]
        elif k == "thermostat" and isinstance(v, str):
            v = THERMO[v]
        elif k == "exchange_algo" and isinstance(v, str):
            v = EXCHANGE[v]
        setattr(p, k, v)
    return p


_lib = None


def lib():
    """Loads libpimdb_gpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpimdb_gpu.so is not built: run __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sigs = {
        "pimdb_ctx_create": (i, [C.POINTER(Params), C.POINTER(vp)]),
        "pimdb_ctx_destroy": (None, [vp]),
        "pimdb_last_error": (C.c_char_p, [vp]),
        "pimdb_set_state": (i, [vp, _DP, _DP]),
        "pimdb_get_state": (i, [vp, _DP, _DP, _DP]),
        "pimdb_forces": (i, [vp]),
        "pimdb_exchange_forward": (i, [vp, _DP, _DP, _DP, _DP, _DP]),
        "pimdb_exchange_adjoint": (i, [vp, _DP, _DP]),
        "pimdb_step": (i, [vp, i, _DP]),
        "pimdb_observables": (i, [vp, _DP]),
        "pimdb_exchange_estimators": (i, [vp, _DP, _DP, _DP]),
        "pimdb_graph_capture": (i, [vp]),
        "pimdb_time_kernels": (i, [vp, i, _DP, _DP, _DP]),
        "pimdb_peak_probe": (i, [_DP, _DP, _DP, C.c_longlong]),
        "pimdb_math_probe": (i, [i, _DP, _DP, C.c_longlong]),
        "pimdb_device_info": (i, [C.POINTER(i), C.POINTER(C.c_longlong)]),
        "pimdb_xplan_create": (i, [i, i, d, d, i, d, C.POINTER(vp)]),
        "pimdb_xplan_destroy": (None, [vp]),
        "pimdb_xplan_forward": (i, [vp, vp, vp, vp, vp, vp, vp]),
        "pimdb_xplan_vb_grad": (i, [vp, vp, vp, vp, vp]),
        "pimdb_exchange_vb_grad_noalloc": (i, [i, i, d, d, i, d, vp, vp, vp, vp, vp]),
        "pimdb_batch_create": (i, [C.POINTER(Params), i, C.POINTER(vp)]),
        "pimdb_batch_destroy": (None, [vp]),
        "pimdb_batch_last_error": (C.c_char_p, [vp]),
        "pimdb_batch_set_state": (i, [vp, i, _DP, _DP]),
        "pimdb_batch_get_state": (i, [vp, i, _DP, _DP, _DP]),
        "pimdb_batch_forces": (i, [vp]),
        "pimdb_batch_step": (i, [vp, i]),
        "pimdb_batch_configure": (i, [vp, i, i]),
        "pimdb_batch_autotune": (i, [vp, i, C.POINTER(i), C.POINTER(i)]),
        "pimdb_batch_time_step": (i, [vp, i, _DP]),
        "pimdb_batch_step_timed": (i, [vp, i, _DP]),
        "pimdb_batch_time_phase": (i, [vp, i, i, _DP]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def _raise(code, msg):
    if code == E_NONFINITE:
        raise NonFiniteError(code, msg)
    raise PimdbError(code, msg)


class Context:
    """One ring-polymer system resident on the GPU (the plain per-kernel form).
    Host arrays are [bead][particle][axis]."""

    def __init__(self, params):
        self.params = params
        h = C.c_void_p()
        rc = lib().pimdb_ctx_create(C.byref(params), C.byref(h))
        if rc != OK:
            _raise(rc, lib().pimdb_last_error(None).decode())
        self.h = h
        self.shape = (params.nbeads, params.natoms, params.ndim)

    def close(self):
        if getattr(self, "h", None):
            lib().pimdb_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc != OK:
            _raise(rc, lib().pimdb_last_error(self.h).decode())

    def set_state(self, x=None, p=None):
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        for a in (x, p):
            if a is not None and a.shape != self.shape:
                raise ValueError("state must have shape %s" % (self.shape,))
        self._chk(lib().pimdb_set_state(self.h, _p(x), _p(p)))

    def state(self):
        x, p, f = (np.zeros(self.shape) for _ in range(3))
        self._chk(lib().pimdb_get_state(self.h, _p(x), _p(p), _p(f)))
        return x, p, f

    def forces(self):
        self._chk(lib().pimdb_forces(self.h))

    def exchange_forward(self, prob=True, E=True, V=True):
        """V_B, and unless switched off V, Vb (V=False drops both), the
        connection probabilities and the packed E_kn triangle.  The factorial
        exchange defines only V_B."""
        N = self.params.natoms
        VB = np.zeros(1)
        Vb = np.zeros(N + 1) if V else None
        V = np.zeros(N + 1) if V else None
        P = np.zeros((N, N)) if prob else None
        Ek = np.zeros(N * (N + 1) // 2) if E else None
        self._chk(lib().pimdb_exchange_forward(self.h, _p(VB), _p(V), _p(Vb), _p(P), _p(Ek)))
        return dict(VB=VB[0], V=V, Vb=Vb, prob=P, E=Ek)

    def exchange_adjoint(self):
        N, D = self.params.natoms, self.params.ndim
        f1 = np.zeros((N, D))
        fP = np.zeros((N, D))
        self._chk(lib().pimdb_exchange_adjoint(self.h, _p(f1), _p(fP)))
        return f1, fP

    def step(self, nsteps=1, noise=None):
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
            if noise.size != nsteps * 2 * int(np.prod(self.shape)):
                raise ValueError("noise must hold nsteps x 2 x P x N x D values")
        self._chk(lib().pimdb_step(self.h, int(nsteps), _p(noise)))

    def observables(self):
        out = np.zeros(11)
        self._chk(lib().pimdb_observables(self.h, _p(out)))
        return out

    def exchange_estimators(self):
        """(primEstimator, distinct probability, longest probability)."""
        v = np.zeros(3)
        self._chk(lib().pimdb_exchange_estimators(self.h, _p(v[0:]), _p(v[1:]), _p(v[2:])))
        return float(v[0]), float(v[1]), float(v[2])

    def graph_capture(self):
        self._chk(lib().pimdb_graph_capture(self.h))

    def time_kernels(self, reps=10, forces=True, exchange=True, step=True):
        vals = [np.zeros(1) if w else None for w in (forces, exchange, step)]
        self._chk(lib().pimdb_time_kernels(self.h, int(reps), *(_p(v) for v in vals)))
        return tuple(None if v is None else float(v[0]) for v in vals)


class Batch:
    """Independent replicas advanced by one persistent kernel per MD step."""

    def __init__(self, params_list):
        arr = (Params * len(params_list))(*params_list)
        h = C.c_void_p()
        rc = lib().pimdb_batch_create(arr, len(params_list), C.byref(h))
        if rc != OK:
            _raise(rc, lib().pimdb_batch_last_error(None).decode())
        self.h = h
        self.params = list(params_list)

    def close(self):
        if getattr(self, "h", None):
            lib().pimdb_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc != OK:
            _raise(rc, lib().pimdb_batch_last_error(self.h).decode())

    def shape(self, r):
        q = self.params[r]
        return (q.nbeads, q.natoms, q.ndim)

    def set_state(self, r, x=None, p=None):
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        for a in (x, p):
            if a is not None and a.shape != self.shape(r):
                raise ValueError("state of replica %d must have shape %s" % (r, self.shape(r)))
        self._chk(lib().pimdb_batch_set_state(self.h, int(r), _p(x), _p(p)))

    def state(self, r):
        x, p, f = (np.zeros(self.shape(r)) for _ in range(3))
        self._chk(lib().pimdb_batch_get_state(self.h, int(r), _p(x), _p(p), _p(f)))
        return x, p, f

    def forces(self):
        self._chk(lib().pimdb_batch_forces(self.h))

    def step(self, nsteps=1):
        self._chk(lib().pimdb_batch_step(self.h, int(nsteps)))

    def configure(self, threads=0, resident=0):
        self._chk(lib().pimdb_batch_configure(self.h, int(threads), int(resident)))

    def autotune(self, trial_steps=3):
        t, r = C.c_int(), C.c_int()
        self._chk(lib().pimdb_batch_autotune(self.h, int(trial_steps), C.byref(t), C.byref(r)))
        return t.value, r.value

    def step_timed(self, nsteps):
        """nsteps steps; returns the device ms between stream events around them."""
        v = np.zeros(1)
        self._chk(lib().pimdb_batch_step_timed(self.h, int(nsteps), _p(v)))
        return float(v[0])

    def time_step(self, reps=5):
        v = np.zeros(1)
        self._chk(lib().pimdb_batch_time_step(self.h, int(reps), _p(v)))
        return float(v[0])

    def time_phase(self, phase, reps=3):
        """phase: 0 step, 1 pair forces, 2 exchange, 3 integrator (ms/launch)."""
        v = np.zeros(1)
        self._chk(lib().pimdb_batch_time_phase(self.h, int(phase), int(reps), _p(v)))
        return float(v[0])


def peak_probe(copy_bytes=1 << 30):
    """Measured FP64 FMA TFLOP/s, FP64 exp Gexp/s and copy GB/s."""
    v = [np.zeros(1) for _ in range(3)]
    rc = lib().pimdb_peak_probe(_p(v[0]), _p(v[1]), _p(v[2]), int(copy_bytes))
    if rc != OK:
        _raise(rc, "pimdb_peak_probe failed")
    return tuple(float(a[0]) for a in v)


def device_info():
    """(compute units, LDS bytes per workgroup) of the current device."""
    cu, lds = C.c_int(), C.c_longlong()
    rc = lib().pimdb_device_info(C.byref(cu), C.byref(lds))
    if rc != OK:
        _raise(rc, "pimdb_device_info failed")
    return cu.value, lds.value


def device_cus():
    return device_info()[0]


def math_probe(fn, x):
    """The device elementary functions evaluated on the GPU at the values x:
    e^x to 1 ulp (fn "exp", the exchange's and integrator's), 1/sqrt(x) to 1
    ulp ("rsqrt", the full-precision form of "rsqrt_pair"), the noise's
    ln x on (0, 1] ("log_unit") and sin / cos of 2 pi x on [0, 1)
    ("sin_2pi", "cos_2pi"), and the pair forces' e^x for x <= 0 ("exp_pair")
    and 1/sqrt(x) ("rsqrt_pair")."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    fns = {"exp": 0, "rsqrt": 1, "log_unit": 2, "sin_2pi": 3, "cos_2pi": 4, "exp_pair": 5, "rsqrt_pair": 6}
    rc = lib().pimdb_math_probe(fns[fn], _p(x), _p(y), int(x.size))
    if rc != OK:
        _raise(rc, "pimdb_math_probe failed")
    return y


class XPlan:
    """Exchange on device arrays (torch tensors' data_ptr): forward V_B and
    -dV_B/dx for beads 1 and P.  Arrays are SoA [axis][particle]."""

    def __init__(self, natoms, ndim, beta_t, spring_k, pbc=False, size=0.0):
        h = C.c_void_p()
        rc = lib().pimdb_xplan_create(natoms, ndim, beta_t, spring_k, int(pbc), size, C.byref(h))
        if rc != OK:
            _raise(rc, "pimdb_xplan_create failed")
        self.h = h
        self.natoms, self.ndim = natoms, ndim

    def close(self):
        if getattr(self, "h", None):
            lib().pimdb_xplan_destroy(self.h)
            self.h = None

    __del__ = close

    def forward(self, x1_ptr, xP_ptr, VB_ptr, V_ptr=None, Vb_ptr=None, stream=None):
        rc = lib().pimdb_xplan_forward(self.h, x1_ptr, xP_ptr, VB_ptr, V_ptr, Vb_ptr, stream)
        if rc != OK:
            _raise(rc, "pimdb_xplan_forward failed")

    def vb_grad(self, x1_ptr, xP_ptr, grad_ptr, stream=None):
        rc = lib().pimdb_xplan_vb_grad(self.h, x1_ptr, xP_ptr, grad_ptr, stream)
        if rc != OK:
            _raise(rc, "pimdb_xplan_vb_grad failed")


def exchange_vb_grad_noalloc(natoms, ndim, beta_t, spring_k, pbc, size, x1_ptr, xP_ptr, VB_ptr=None,
                             grad_ptr=None, stream=None):
    """V_B and -dV_B/dx of beads 1 and P on device arrays (SoA [axis][particle];
    grad [2][D][N]) with no device allocation: the library's fixed per-GPU
    workspace, shared by every call on a GPU (overlapping calls race)."""
    rc = lib().pimdb_exchange_vb_grad_noalloc(int(natoms), int(ndim), float(beta_t), float(spring_k), int(pbc),
                                              float(size), x1_ptr, xP_ptr, VB_ptr, grad_ptr, stream)
    if rc != OK:
        _raise(rc, lib().pimdb_last_error(None).decode())
