"""The C-ABI library loads and exports every symbol include/pimdb_gpu.h
declares; host-side argument checks work without a GPU (no compute calls)."""
import ctypes as C
import os
import re

import pytest

import pimdb_amd as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pimdb_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(G.LIB_PATH)
    names = declared("pimdb_gpu.h")
    assert "pimdb_ctx_create" in names and "pimdb_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_parameters_rejected_before_any_device_work():
    bad = [dict(natoms=0, nbeads=4), dict(natoms=4, nbeads=0), dict(natoms=4, nbeads=4, ndim=4),
           dict(natoms=4, nbeads=4, bosonic=1, propagator="normal_modes"),
           dict(natoms=4, nbeads=4, thermostat="none", nmthermostat=1)]
    for kw in bad:
        base = dict(mass=1.0, temperature=1e-5, dt=1.0, gamma=0.01, size=10.0)
        base.update(kw)
        with pytest.raises(G.PimdbError) as ei:
            G.Context(G.make_params(**base))
        assert ei.value.code == G.E_ARG


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_gpu_is_an_error_not_a_fallback():
    with pytest.raises(G.PimdbError) as ei:
        G.Context(G.make_params(natoms=4, nbeads=4, mass=1.0, temperature=1e-5, dt=1.0, size=10.0))
    assert ei.value.code == G.E_HIP
