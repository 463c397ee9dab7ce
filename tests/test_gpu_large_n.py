"""Rings larger than one workgroup's LDS on the single-system form, through
the C ABI:

* the exchange recursions with their workspace in device memory (the LDS
  workspace holds (2D+4)N + 4194 doubles: N above 1628 at D=3 and 2035 at
  D=2 on a 160 KB workgroup), through the context and through the exchange
  plan;
* the pair forces and pair energies reading their slice in place (the
  staged slice holds 3N + 2880 doubles at D=3: N above 5866);
* batched replicas whose exchange layout exceeds the workgroup's LDS share
  (about 15N + 3000 doubles at D=3: N above about 470 with two 512-thread
  workgroups per CU, 1080 with one 1024-thread workgroup), whose exchange
  then runs on a workspace in device memory.

The exchange is judged against the exact exchange of the same double inputs
(oracle.exchange_quad, __float128), as the N=1024 rings are: these rings are
conditioned so that the reference-order oracle is itself 4e-10 (D=3, N=1800)
and 1e-9 (D=2, N=2200) from the exact forces, since a direct-link
probability 1 - exp(-x), x ~ 0, carries an absolute error of about
beta~ eps |V| in any double evaluation.  The GPU must be at least as close
to the exact values as the oracle.  The pair forces carry no exchange term
and are compared with the oracle at 1e-10 per bead."""
import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, KELVIN, MEV, aziz_config, per_bead_err, rel_err

pytestmark = pytest.mark.gpu


def thermal_ring(N, P, seed, ndim=3, density=0.02186):
    """He-4 at 1.5 K (PBC, Aziz, cutoff L/2) at the given number density
    (per A^ndim): seeded positions replicated over the beads, then displaced
    by 0.3 bohr."""
    kw, x = aziz_config(N, P, 1.5, replicate=True, seed=seed, ndim=ndim, density_A3=density)
    x = x + np.random.default_rng(seed + 1).normal(0.0, 0.3, size=x.shape)
    return kw, x


def ring_args(kw, x):
    d = O.derived(O.make_params(**kw))
    P = kw["nbeads"]
    return (x[0], x[1 % P], x[(P - 2) % P], x[P - 1], d["beta_t"], d["k"], kw["pbc"], kw["size"]), d


def context_exchange(kw, x):
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x)
    g = ctx.exchange_forward()
    f = ctx.exchange_adjoint()
    ctx.close()
    return g, f


@pytest.mark.parametrize("N,P,ndim,density", [(1800, 8, 3, 0.02186), (2200, 6, 2, 0.08)])
def test_exchange_device_workspace(N, P, ndim, density):
    kw, x = thermal_ring(N, P, 100 + N, ndim, density)
    g, f = context_exchange(kw, x)
    args, _ = ring_args(kw, x)
    o = O.exchange(*args)
    q = O.exchange_quad(*args)
    assert rel_err(g["E"], o["E"]) < 1e-12
    assert abs(g["VB"] - g["V"][N]) == 0.0
    Pm = g["prob"]
    assert np.allclose(Pm.sum(axis=1), 1.0, atol=1e-10)
    assert np.allclose(Pm.sum(axis=0), 1.0, atol=1e-10)
    errs = {}
    for name, gv, ov, qv, kind in (("V", g["V"], o["V"], q["V"], "rel"), ("Vb", g["Vb"], o["Vb"], q["Vb"], "rel"),
                                   ("prob", g["prob"], o["prob"], q["prob"], "abs"),
                                   ("f_first", f[0], o["f_first"], q["f_first"], "rel"),
                                   ("f_last", f[1], o["f_last"], q["f_last"], "rel")):
        e = (lambda a: float(np.max(np.abs(a - qv)))) if kind == "abs" else (lambda a: rel_err(a, qv))
        errs[name] = (e(gv), e(ov))
        print("N=%d D=%d %s: GPU - exact %.2e, oracle - exact %.2e" % (N, ndim, name, errs[name][0], errs[name][1]),
              flush=True)
    # V and Vb within a few ulps (1e-14, the bar of the N=1024 rings), the
    # rest at least as close as the oracle (5% for rounding ties)
    for name, (eg, eo) in errs.items():
        assert eg <= (max(1e-14, eo) if name in ("V", "Vb") else 1.05 * eo), (name, eg, eo)


class DeviceArrays:
    """Device buffers from the HIP runtime the library itself links (the
    plan's arguments are device pointers)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, shape):
        a = np.zeros(shape)
        assert self.hip.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0   # device to host
        return a

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def test_exchange_plan_device_workspace():
    """The exchange plan (pimdb_xplan_*) at N=1800 against the context on the
    same ring: V_B, and -dV_B/dx as the context's exterior forces less the
    interior springs k (x2 - x1) and k (x_{P-1} - x_P) (minimum image)."""
    N, P = 1800, 8
    kw, x = thermal_ring(N, P, 1900)
    g, f = context_exchange(kw, x)
    _, d = ring_args(kw, x)
    mem = DeviceArrays()
    x1 = mem.put(x[0].T)   # [D][N]
    xP = mem.put(x[P - 1].T)
    vb = mem.put(np.zeros(1))
    grad = mem.put(np.zeros((2, 3, N)))
    plan = G.XPlan(N, 3, d["beta_t"], d["k"], 1, kw["size"])
    plan.forward(x1, xP, vb, None, None, None)
    plan.vb_grad(x1, xP, grad, None)
    plan.close()
    VB = mem.get(vb, (1,))[0]
    gr = mem.get(grad, (2, 3, N))
    mem.free()
    L = kw["size"]

    def img(t):
        return t - L * np.rint(t / L)

    g1 = f[0] - d["k"] * img(x[1] - x[0])
    gP = f[1] - d["k"] * img(x[P - 2] - x[P - 1])
    print("plan: VB %.2e grad1 %.2e gradP %.2e" % (abs(VB - g["VB"]) / abs(g["VB"]), rel_err(gr[0].T, g1),
                                                   rel_err(gr[1].T, gP)), flush=True)
    assert abs(VB - g["VB"]) <= 1e-14 * abs(g["VB"])
    assert rel_err(gr[0].T, g1) < 1e-12
    assert rel_err(gr[1].T, gP) < 1e-12


def test_pair_forces_and_energies_slice_in_place():
    """N=6000, P=4: the pair forces read in place, on the interior beads
    (pairs and springs) and on beads 1 and P (the total less the exchange
    adjoint: the pair forces alone), and the observables (pair energies and
    virial read in place) against the oracle."""
    N, P = 6000, 4
    kw, x = thermal_ring(N, P, 6000)
    p = np.random.default_rng(6002).normal(0.0, 0.3, size=x.shape)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    ctx.forces()
    _, _, fg = ctx.state()
    f1, fP = ctx.exchange_adjoint()
    og = ctx.observables()
    ctx.close()
    prm = O.make_params(**kw)
    fo = O.update_forces(prm, x)
    errs = [rel_err(fg[b], fo[b]) for b in (1, 2)]
    for b, fx in ((0, f1), (P - 1, fP)):
        errs.append(rel_err(fg[b] - fx, O.physical_forces(prm, x[b])))
    sim = O.Sim(prm)
    sim.set_state(x, p)
    oo = sim.observables()
    print("N=6000 pair forces, beads 2, 3, 1, P:", ["%.2e" % e for e in errs], flush=True)
    for e in errs:
        assert e < 1e-10, errs
    names = ["kinetic", "potential", "ext_pot", "int_pot", "virial", "temperature", "cl_kinetic", "cl_spring",
             "prob_dist", "prob_all"]
    for j, name in enumerate(names):
        print("  %s GPU %.15e oracle %.15e" % (name, og[j], oo[j]), flush=True)
    for j, name in enumerate(names[:8]):
        assert abs(og[j] - oo[j]) <= 1e-9 * max(abs(oo[j]), 1e-12), (name, og[j], oo[j])
    for j in (8, 9):
        assert abs(og[j] - oo[j]) <= 1e-10 + 1e-9 * abs(oo[j]), (names[j], og[j], oo[j])


@pytest.mark.parametrize("shape", ["512x2", "1024x1"])
def test_batch_exchange_device_workspace(shape):
    """N=256 (LDS) beside a replica whose exchange runs on device memory
    (N=600 on two 512-thread workgroups per CU, N=1200 on one 1024-thread
    workgroup), P=4, Aziz with PBC: every replica's forces against the oracle
    per bead at 1e-10, then 10 device-noise steps against the single-system
    context (x to 1e-10, p and forces to 1e-9, as the batched tests of the
    other paths)."""
    cus = G.device_cus()
    threads, resident, Nbig = (512, 2 * cus, 600) if shape == "512x2" else (1024, cus, 1200)
    reps = []
    for r, N in enumerate([256, Nbig]):
        kw, x = aziz_config(N, 4, 1.5, replicate=False, seed=700 + r)
        kw["seed"] = 31 + r
        p = np.random.default_rng(70 + r).normal(0, 0.3, size=x.shape)
        reps.append((kw, x, p))
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    b.configure(threads, resident)
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    errs = []
    for r, (kw, x, p) in enumerate(reps):
        errs.append(per_bead_err(b.state(r)[2], O.update_forces(O.make_params(**kw), x)))
    print(shape, "per-replica force errors", errs, flush=True)
    assert max(errs) < 1e-10, errs
    ctxs = []
    for kw, x, p in reps:
        c = G.Context(G.make_params(**kw))
        c.set_state(x, p)
        c.forces()
        ctxs.append(c)
    b.step(10)
    for r, c in enumerate(ctxs):
        c.step(10)
        xb, pb, fb = b.state(r)
        xc, pc, fc = c.state()
        print(shape, r, "x %.2e p %.2e f %.2e" % (rel_err(xb, xc), rel_err(pb, pc), per_bead_err(fb, fc)), flush=True)
        assert rel_err(xb, xc) < 1e-10, r
        assert rel_err(pb, pc) < 1e-9, r
        assert per_bead_err(fb, fc) < 1e-9, r
        c.close()
    b.close()


def lattice_dipole_ring(N, P, T_kelvin, seed, nmthermostat):
    """D=2 He-4 with a 3 meV trap and dipole pairs (all pairs), particles on a
    jittered 4-bohr lattice (at least 3 bohr apart) and beads 0.3 bohr about
    them, as the batched form's config-4 replicas."""
    rng = np.random.default_rng(seed)
    kw = dict(natoms=N, nbeads=P, ndim=2, mass=4.0026 * AMU, temperature=T_kelvin * KELVIN, dt=20.0,
              gamma=1.0 / 2000.0, bosonic=1, pbc=0, size=60.0, ext_kind="harmonic", ext_omega=3.0 * MEV,
              int_kind="dipole", int_strength=1.3, cutoff=-1.0, propagator="cartesian", thermostat="langevin",
              nmthermostat=nmthermostat, fixcom=1, seed=seed)
    side = int(np.ceil(np.sqrt(N)))
    grid = np.array([(a, c) for a in range(side) for c in range(side)][:N], dtype=float) * 4.0
    centres = grid - grid.mean(axis=0) + rng.uniform(-0.5, 0.5, size=(N, 2))
    x = np.repeat(centres[None], P, axis=0) + rng.normal(0, 0.3, size=(P, N, 2))
    p = rng.normal(0, 0.5, size=(P, N, 2))
    return kw, x, p


def test_batch_exchange_device_workspace_2d_dipole():
    """D=2 dipole replicas on two 512-thread workgroups per CU: N=64 (LDS)
    beside N=700 (exchange workspace in device memory, about 11.6k doubles
    against 10.2k), one with the normal-mode thermostat: forces against the
    oracle per bead at 1e-10, then 10 device-noise steps against the
    single-system context."""
    cus = G.device_cus()
    reps = [lattice_dipole_ring(64, 8, 2.0, 81, 0), lattice_dipole_ring(700, 4, 2.0, 82, 1)]
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    b.configure(512, 2 * cus)
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    errs = [per_bead_err(b.state(r)[2], O.update_forces(O.make_params(**kw), x)) for r, (kw, x, p) in enumerate(reps)]
    print("2d dipole per-replica force errors", errs, flush=True)
    assert max(errs) < 1e-10, errs
    ctxs = []
    for kw, x, p in reps:
        c = G.Context(G.make_params(**kw))
        c.set_state(x, p)
        c.forces()
        ctxs.append(c)
    b.step(10)
    for r, c in enumerate(ctxs):
        c.step(10)
        xb, pb, fb = b.state(r)
        xc, pc, fc = c.state()
        print("2d dipole", r, "x %.2e p %.2e f %.2e" % (rel_err(xb, xc), rel_err(pb, pc), per_bead_err(fb, fc)),
              flush=True)
        assert rel_err(xb, xc) < 1e-10, r
        assert rel_err(pb, pc) < 1e-9, r
        assert per_bead_err(fb, fc) < 1e-9, r
        c.close()
    b.close()


def test_batch_device_workspace_autotune_and_phase_timing():
    """The autotuner over every launch shape that fits (each re-plans the LDS
    and re-allocates the device-memory exchange workspaces) and the phase
    timer (the exchange phase through the kernel that reaches them) on a
    batch holding an N=600 replica beside an N=256 one: the state it
    restores gives the oracle's forces per bead at 1e-10."""
    reps = []
    for r, N in enumerate([256, 600]):
        kw, x = aziz_config(N, 4, 1.5, replicate=False, seed=900 + r)
        kw["seed"] = 41 + r
        p = np.random.default_rng(90 + r).normal(0, 0.3, size=x.shape)
        reps.append((kw, x, p))
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    threads, resident = b.autotune(trial_steps=2)
    print("autotune picked", threads, resident, flush=True)
    for phase in range(4):
        assert b.time_phase(phase, 2) > 0.0
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    errs = [per_bead_err(b.state(r)[2], O.update_forces(O.make_params(**kw), x)) for r, (kw, x, p) in enumerate(reps)]
    print("after autotune: per-replica force errors", errs, flush=True)
    assert max(errs) < 1e-10, errs
    b.close()
