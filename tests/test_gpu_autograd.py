"""torch.autograd harness over the exchange kernels: forward V_B and the
hand-written adjoint (-dV_B/dx of beads 1 and P) against the oracle and
against central finite differences of the forward kernel itself.

Cases follow two benchmark configurations:
  - N=64 bosons, P=64, T=2 K, He-4 mass, positions uniform in a 2 nm box
    (independent beads, as the reference's `random` init), no PBC;
    FD with h = 1e-5 bohr on every coordinate of beads 1 and P;
  - N=1024 He-4, P=128, T=1.2 K, PBC at 0.02186 A^-3 (L = 68.12 bohr), bead
    P displaced from bead 1 by a thermal Gaussian; FD on a seeded subset of
    16 particles (both beads, all axes).
The oracle's exterior forces include the interior links (x2 - x1, xPm1 - xP);
passing x2 = x1 and xPm1 = xP removes them, leaving -dV_B/dx exactly."""
import numpy as np
import pytest

import oracle as O
from helpers import AMU, ANGSTROM, KELVIN

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _ag():
    from pimdb_amd import autograd as AG
    return AG


def exchange_constants(P, T_kelvin, mass):
    beta = 1.0 / (T_kelvin * KELVIN)
    beta_t = beta / P                      # IPI convention, bosonic_exchange_base.cpp:16-18
    k = mass * (P / beta) ** 2             # simulation.cpp:41-51
    return beta_t, k


def oracle_grad(x1, xP, beta_t, k, pbc, L):
    ex = O.exchange(x1, x1, xP, xP, beta_t, k, pbc, L)
    return ex["V"][-1], -ex["f_first"], -ex["f_last"]


def fd_grad(fun, x, idx, h):
    """Central differences of fun at x along the listed (row, axis) entries."""
    g = np.zeros(len(idx))
    for n, (i, a) in enumerate(idx):
        xp = x.clone()
        xm = x.clone()
        xp[i, a] += h
        xm[i, a] -= h
        g[n] = (float(fun(xp)) - float(fun(xm))) / (2 * h)
    return g


def run_case(N, D, x1_np, xP_np, beta_t, k, pbc, L, fd_rows, h, tol_oracle):
    AG = _ag()
    dev = torch.device("cuda:0")
    plan = AG.make_plan(N, D, beta_t, k, pbc, L)
    x1 = torch.tensor(x1_np, dtype=torch.float64, device=dev, requires_grad=True)
    xP = torch.tensor(xP_np, dtype=torch.float64, device=dev, requires_grad=True)
    vb = AG.exchange_potential(x1, xP, plan)
    vb.backward()
    g1 = x1.grad.cpu().numpy()
    gP = xP.grad.cpu().numpy()

    vb_ref, g1_ref, gP_ref = oracle_grad(x1_np, xP_np, beta_t, k, pbc, L)
    vbv = float(vb.detach())
    assert abs(vbv - vb_ref) <= 1e-10 * abs(vb_ref), (vbv, vb_ref)
    for g, gr in ((g1, g1_ref), (gP, gP_ref)):
        err = np.max(np.abs(g - gr)) / np.max(np.abs(gr))
        assert err <= tol_oracle, err

    # finite differences of the forward kernel, both beads
    idx = [(i, a) for i in fd_rows for a in range(D)]
    with torch.no_grad():
        x1d, xPd = x1.detach(), xP.detach()
        fd1 = fd_grad(lambda y: AG.exchange_potential(y, xPd, plan), x1d, idx, h)
        fdP = fd_grad(lambda y: AG.exchange_potential(x1d, y, plan), xPd, idx, h)
    an1 = np.array([g1[i, a] for i, a in idx])
    anP = np.array([gP[i, a] for i, a in idx])
    scale = max(np.max(np.abs(an1)), np.max(np.abs(anP)))
    assert np.max(np.abs(fd1 - an1)) <= 1e-6 * scale, np.max(np.abs(fd1 - an1)) / scale
    assert np.max(np.abs(fdP - anP)) <= 1e-6 * scale, np.max(np.abs(fdP - anP)) / scale
    plan.close()


def test_autograd_n64_p64_box():
    N, D, P = 64, 3, 64
    mass = 4.0026 * AMU
    beta_t, k = exchange_constants(P, 2.0, mass)
    L = 20.0 * ANGSTROM
    rng = np.random.default_rng(2024)
    x1 = rng.uniform(-0.5 * L, 0.5 * L, size=(N, D))
    xP = rng.uniform(-0.5 * L, 0.5 * L, size=(N, D))
    run_case(N, D, x1, xP, beta_t, k, False, 0.0, list(range(N)), 1e-5, 1e-10)


def test_autograd_n1024_p128_aziz_subset():
    N, D, P = 1024, 3, 128
    mass = 4.0026 * AMU
    beta_t, k = exchange_constants(P, 1.2, mass)
    L = (N / 0.02186) ** (1.0 / 3.0) * ANGSTROM
    rng = np.random.default_rng(77)
    x1 = rng.uniform(-0.5 * L, 0.5 * L, size=(N, D))
    sigma = np.sqrt(1.0 / (beta_t * k))    # thermal width of one ring link
    xP = x1 + rng.normal(0.0, sigma, size=(N, D))
    rows = sorted(rng.choice(N, size=16, replace=False).tolist())
    # thermal N=1024 exchange forces meet 1e-9 against the oracle (conditioning)
    run_case(N, D, x1, xP, beta_t, k, True, L, rows, 1e-5, 1e-9)


def test_autograd_gradcheck_small():
    AG = _ag()
    N, D = 5, 3
    beta_t, k = exchange_constants(8, 1.0, 4.0026 * AMU)
    rng = np.random.default_rng(3)
    x1 = torch.tensor(rng.normal(0, 1.5, size=(N, D)), dtype=torch.float64, device="cuda:0", requires_grad=True)
    xP = torch.tensor(rng.normal(0, 1.5, size=(N, D)), dtype=torch.float64, device="cuda:0", requires_grad=True)
    plan = AG.make_plan(N, D, beta_t, k)
    assert torch.autograd.gradcheck(lambda a, b: AG.exchange_potential(a, b, plan), (x1, xP),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)
    plan.close()


def test_autograd_rejects_float32():
    AG = _ag()
    plan = AG.make_plan(4, 3, 1.0, 1.0)
    x = torch.zeros((4, 3), dtype=torch.float32, device="cuda:0")
    with pytest.raises(TypeError):
        AG.exchange_potential(x, x, plan)
    plan.close()
