"""ctypes bindings to libpimdb_host.so (include/pimdb_host.h): the host C++
side of the GPU path -- configuration files, units, RanMars streams, initial
conditions and the run driver."""
import ctypes as C
import os

import numpy as np

from . import OK, Params, PimdbError, NonFiniteError, E_NONFINITE, lib as gpu_lib

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB_PATH = os.path.join(_PKG, "libpimdb_host.so")
DRIVER_PATH = os.path.join(_PKG, "pimdb_gpu")
_DP = C.POINTER(C.c_double)
_hl = None


def hlib():
    """Loads libpimdb_host.so (and the GPU library it links); raises if it is
    not built."""
    global _hl
    if _hl is not None:
        return _hl
    if not os.path.exists(HOST_LIB_PATH):
        raise RuntimeError("libpimdb_host.so is not built: run __graft_entry__.build()")
    gpu_lib()
    L = C.CDLL(HOST_LIB_PATH)
    vp, i, d, cp, lg = C.c_void_p, C.c_int, C.c_double, C.c_char_p, C.c_long
    sigs = {
        "pimdb_host_to_internal": (i, [cp, cp, d, _DP, cp, i]),
        "pimdb_host_ranmars_new": (vp, [i]),
        "pimdb_host_ranmars_uniform": (d, [vp]),
        "pimdb_host_ranmars_gaussian": (d, [vp]),
        "pimdb_host_ranmars_free": (None, [vp]),
        "pimdb_host_load_config": (i, [cp, C.POINTER(Params), C.POINTER(lg), C.POINTER(lg), cp, i]),
        "pimdb_host_initial_state": (i, [cp, _DP, _DP, cp, i]),
        "pimdb_host_draw_noise": (i, [cp, lg, _DP, cp, i]),
        "pimdb_host_run": (i, [cp, lg, i, cp, cp, i]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _hl = L
    return L


def _call(fn, *args):
    err = C.create_string_buffer(512)
    rc = fn(*args, err, len(err))
    if rc != OK:
        msg = err.value.decode(errors="replace")
        if rc == E_NONFINITE:
            raise NonFiniteError(rc, msg)
        raise PimdbError(rc, msg)


def to_internal(family, unit, number):
    out = C.c_double()
    _call(hlib().pimdb_host_to_internal, family.encode(), unit.encode(), float(number), C.byref(out))
    return out.value


class RanMars:
    def __init__(self, seed):
        self.h = hlib().pimdb_host_ranmars_new(int(seed))
        if not self.h:
            raise ValueError("Invalid seed for Marsaglia random # generator")

    def uniform(self):
        return hlib().pimdb_host_ranmars_uniform(self.h)

    def gaussian(self):
        return hlib().pimdb_host_ranmars_gaussian(self.h)

    def close(self):
        if getattr(self, "h", None):
            hlib().pimdb_host_ranmars_free(self.h)
            self.h = None

    __del__ = close


def load_config(path):
    """(pimdb params, steps, sfreq) of a config.ini."""
    p = Params()
    steps, sfreq = C.c_long(), C.c_long()
    _call(hlib().pimdb_host_load_config, os.fsencode(path), C.byref(p), C.byref(steps), C.byref(sfreq))
    return p, steps.value, sfreq.value


def initial_state(path):
    p, _, _ = load_config(path)
    shape = (p.nbeads, p.natoms, p.ndim)
    x, m = np.zeros(shape), np.zeros(shape)
    _call(hlib().pimdb_host_initial_state, os.fsencode(path), x.ctypes.data_as(_DP), m.ctypes.data_as(_DP))
    return x, m


def draw_noise(path, nsteps):
    p, _, _ = load_config(path)
    out = np.zeros((nsteps, 2, p.nbeads, p.natoms, p.ndim))
    _call(hlib().pimdb_host_draw_noise, os.fsencode(path), int(nsteps), out.ctypes.data_as(_DP))
    return out


def run(path, steps=-1, device_noise=False, output_dir="output", factorial=False):
    """Simulation::run on the GPU (see pimdb_host_run); factorial selects the
    factorial bosonic exchange."""
    flags = (1 if device_noise else 0) | (2 if factorial else 0)
    _call(hlib().pimdb_host_run, os.fsencode(path), int(steps), flags, os.fsencode(output_dir))
