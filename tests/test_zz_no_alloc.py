"""The exchange entry point that allocates nothing
(pimdb_exchange_vb_grad_noalloc) against the plan-based entry points it
varies (pimdb_xplan_forward + pimdb_xplan_vb_grad): the same kernels, so V_B
and -dV_B/dx must agree bit for bit, on two calls in a row (the second with a
smaller ring, which reuses the fixed workspace the first left behind), and
the size limit is an argument error."""
import numpy as np
import pytest

import pimdb_amd as G
from helpers import AMU, ANGSTROM, KELVIN

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def ring(N, D, seed, pbc):
    """Beads 1 and P of a He-4 ring at 1.5 K, P=64: bead P a thermal
    displacement of bead 1; PBC at 0.02186 A^-3 when asked."""
    P, T = 64, 1.5
    mass = 4.0026 * AMU
    beta_t = 1.0 / (T * KELVIN) / P
    k = mass * (P * T * KELVIN) ** 2
    L = (N / 0.02186) ** (1.0 / 3.0) * ANGSTROM if pbc else 0.0
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(-0.5 * (L or 20.0), 0.5 * (L or 20.0), size=(D, N))
    xP = x1 + rng.normal(0.0, 1.0 / np.sqrt(beta_t * k), size=(D, N))
    return x1, xP, beta_t, k, L


def by_plan(N, D, x1, xP, beta_t, k, pbc, L, dev):
    plan = G.XPlan(N, D, beta_t, k, pbc, L)
    vb = torch.zeros(1, dtype=torch.float64, device=dev)
    grad = torch.zeros((2, D, N), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    plan.forward(x1.data_ptr(), xP.data_ptr(), vb.data_ptr(), None, None, stream)
    plan.vb_grad(x1.data_ptr(), xP.data_ptr(), grad.data_ptr(), stream)
    plan.close()
    return vb.cpu().numpy(), grad.cpu().numpy()


def by_noalloc(N, D, x1, xP, beta_t, k, pbc, L, dev, with_grad=True):
    vb = torch.zeros(1, dtype=torch.float64, device=dev)
    grad = torch.zeros((2, D, N), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    G.exchange_vb_grad_noalloc(N, D, beta_t, k, pbc, L, x1.data_ptr(), xP.data_ptr(), vb.data_ptr(),
                               grad.data_ptr() if with_grad else None, stream)
    return vb.cpu().numpy(), grad.cpu().numpy()


def test_noalloc_equals_plan_on_two_calls():
    dev = torch.device("cuda:0")
    for N, D, pbc, seed in [(300, 3, 1, 5), (72, 2, 0, 6)]:
        x1n, xPn, beta_t, k, L = ring(N, D, seed, pbc)
        x1 = torch.tensor(x1n, dtype=torch.float64, device=dev)
        xP = torch.tensor(xPn, dtype=torch.float64, device=dev)
        vb_p, g_p = by_plan(N, D, x1, xP, beta_t, k, pbc, L, dev)
        vb_n, g_n = by_noalloc(N, D, x1, xP, beta_t, k, pbc, L, dev)
        assert np.isfinite(vb_p).all() and np.abs(g_p).max() > 0.0
        assert np.array_equal(vb_n, vb_p), (N, vb_n, vb_p)
        assert np.array_equal(g_n, g_p), (N, float(np.abs(g_n - g_p).max()))


def test_noalloc_forward_only_and_limit():
    dev = torch.device("cuda:0")
    N, D = 40, 3
    x1n, xPn, beta_t, k, L = ring(N, D, 9, 1)
    x1 = torch.tensor(x1n, dtype=torch.float64, device=dev)
    xP = torch.tensor(xPn, dtype=torch.float64, device=dev)
    vb_p, _ = by_plan(N, D, x1, xP, beta_t, k, 1, L, dev)
    vb_n, g_n = by_noalloc(N, D, x1, xP, beta_t, k, 1, L, dev, with_grad=False)
    assert np.array_equal(vb_n, vb_p)
    assert not np.any(g_n)   # grad not asked for: left untouched
    big = 4096
    z = torch.zeros((D, big), dtype=torch.float64, device=dev)
    with pytest.raises(G.PimdbError) as ei:
        G.exchange_vb_grad_noalloc(big, D, beta_t, k, 0, 0.0, z.data_ptr(), z.data_ptr())
    assert ei.value.code == G.E_ARG
    assert "maximum" in str(ei.value)
