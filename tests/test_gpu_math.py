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


def test_exp_pair_relative_bound():
    """The Aziz force's e^x (x <= 0): within 5e-14 relative of the correctly
    rounded value (normal results), 2e-14 on the repulsion's range x >= -51;
    0 below -746."""
    rng = np.random.default_rng(13)
    x = np.concatenate([rng.uniform(-51.0, 0.0, 1500), rng.uniform(-708.0, -51.0, 800),
                        -np.arange(0, 60) * np.log(2.0) / 2, [0.0, -1e-300, -5e-324, -0.5 * np.log(2.0)]])
    y = G.math_probe("exp_pair", x)
    ref = correctly_rounded(mp.exp, x)
    rel = np.abs(y - ref) / ref
    near = x >= -51.0
    assert rel[near].max() <= 2e-14, (x[near][np.argmax(rel[near])], rel[near].max())
    assert rel.max() <= 5e-14, (x[np.argmax(rel)], rel.max())
    assert np.all(G.math_probe("exp_pair", np.array([-746.5, -750.0, -1e4, -1e300, -np.inf])) == 0.0)


def test_rsqrt_pair_relative_bound():
    """The pair forces' 1/sqrt(x): within 2e-14 relative."""
    rng = np.random.default_rng(14)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), rng.uniform(1.0, 2000.0, 1500),
                        [1.0, 2.0, 4.0, 0.5, 1e-300, 1e300]])
    y = G.math_probe("rsqrt_pair", x)
    ref = correctly_rounded(lambda v: 1 / mp.sqrt(v), x)
    rel = np.abs(y - ref) / ref
    assert rel.max() <= 2e-14, (x[np.argmax(rel)], rel.max())


def test_rsqrt_within_one_ulp():
    rng = np.random.default_rng(12)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), rng.uniform(1.0, 500.0, 1500),
                        [1.0, 4.0, 0.25, 2.0, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0)]])
    y = G.math_probe("rsqrt", x)
    ref = correctly_rounded(lambda v: 1 / mp.sqrt(v), x)
    err = ulp_err(y, ref)
    assert err.max() <= 1.0, (x[np.argmax(err)], err.max())


def test_log_unit_within_two_ulp():
    """ln u on (0, 1], the Box-Muller radius of the device noise."""
    rng = np.random.default_rng(13)
    x = np.concatenate([rng.uniform(0.0, 1.0, 1500), 2.0 ** -rng.uniform(0, 53, 1000),
                        [1.0, np.nextafter(1.0, 0.0), 0.5, np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 1.0),
                         2.0 ** -53, 1.1102230246251565e-16 * 3]])
    x = x[x > 0.0]
    y = G.math_probe("log_unit", x)
    ref = correctly_rounded(mp.log, x)
    err = ulp_err(y, ref)
    assert err[x != 1.0].max() <= 2.0, (x[np.argmax(err)], err.max())
    assert y[x == 1.0].max() == 0.0


def test_sincos_2pi_absolute():
    """sin and cos of 2 pi u on [0, 1), the Box-Muller angle: within 2e-16 of
    the exact value (the argument's own rounding bounds the rest)."""
    rng = np.random.default_rng(14)
    x = np.concatenate([rng.uniform(0.0, 1.0, 2000), np.arange(0, 64) / 64.0,
                        [np.nextafter(1.0, 0.0), 2.0 ** -53, 0.125, 0.375, 0.625, 0.875]])
    mp.mp.dps = 40
    ref_s = np.array([float(mp.sin(2 * mp.pi * mp.mpf(float(v)))) for v in x])
    ref_c = np.array([float(mp.cos(2 * mp.pi * mp.mpf(float(v)))) for v in x])
    s = G.math_probe("sin_2pi", x)
    c = G.math_probe("cos_2pi", x)
    assert np.max(np.abs(s - ref_s)) <= 2e-16
    assert np.max(np.abs(c - ref_c)) <= 2e-16
    assert np.max(np.abs(s * s + c * c - 1.0)) <= 4e-16
