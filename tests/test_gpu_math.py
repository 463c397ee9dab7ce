"""The pair kernels' own elementary functions on the GPU against correctly
rounded values (mpmath at 40 digits): e^x (Cody-Waite reduction and a
degree-11 minimax polynomial) and 1/sqrt(x) (hardware estimate and one
third-order correction).  The reference calls the C library's exp and sqrt
(aziz.cpp:56-106), which are correctly rounded or within 1 ulp."""
import mpmath as mp
import numpy as np
import pytest

import pimdb_amd as G

pytestmark = pytest.mark.gpu


def ulp_err(y, ref):
    """|y - ref| in units of the last place of ref (subnormal spacing below
    the normal range)."""
    return np.abs(y - ref) / np.spacing(np.abs(ref))


def correctly_rounded(f, x):
    mp.mp.dps = 40
    return np.array([float(f(mp.mpf(float(v)))) for v in x])


def test_exp_within_one_ulp():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-708.0, 0.0, 1500),      # the pair kernels' range
                        rng.uniform(-1.0, 1.0, 500),
                        rng.uniform(0.0, 709.0, 200),
                        -np.arange(0, 60) * np.log(2.0) / 2,  # reduction boundaries
                        [0.0, -1e-300, -5e-324, 1e-17, -0.5 * np.log(2.0), 0.5 * np.log(2.0)]])
    y = G.math_probe("exp", x)
    ref = correctly_rounded(mp.exp, x)
    err = ulp_err(y, ref)
    assert err.max() <= 1.0, (x[np.argmax(err)], err.max())


def test_exp_subnormal_and_underflow():
    x = np.concatenate([np.linspace(-745.1, -708.5, 400), [-746.0, -750.0, -1e4, -1e300, -np.inf]])
    y = G.math_probe("exp", x)
    ref = correctly_rounded(mp.exp, x[:400])
    # results are subnormal: at most one unit of the subnormal spacing apart
    assert np.max(np.abs(y[:400] - ref) / 5e-324) <= 1.0
    assert np.all(y[400:] == 0.0)


def test_rsqrt_within_one_ulp():
    rng = np.random.default_rng(12)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), rng.uniform(1.0, 500.0, 1500),
                        [1.0, 4.0, 0.25, 2.0, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0)]])
    y = G.math_probe("rsqrt", x)
    ref = correctly_rounded(lambda v: 1 / mp.sqrt(v), x)
    err = ulp_err(y, ref)
    assert err.max() <= 1.0, (x[np.argmax(err)], err.max())
