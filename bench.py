"""Benchmark: MD steps/s of the bosonic PIMD hot path on one MI355X.

Workload (one "step" = one OBABO MD step of the full ring polymer): N=256
He-4 bosons, P=64 beads, D=3, T=1.5 K, Aziz HFDHE2 pairs, PBC cube at
0.02186 A^-3 (L = 42.914 bohr), cutoff L/2 with minimum image, no external
potential, dt = 0.5 fs, Langevin thermostat with normal-mode coupling,
fixcom, FP64.  Positions: seeded uniform with a 2.2 A minimum-distance
rejection (raw uniform He at liquid density overlaps cores), replicated over
beads; momenta Maxwell-Boltzmann at beta/P.  Device (Philox) noise.

Forms measured:
  plain    one system, one kernel per stage, captured as a hipGraph
  batched  R independent replicas of the same workload in one persistent
           kernel per MD step (every replica starts from the same x and p and
           draws its own noise stream)

Also reported: the batched rate after the workload's 2000-step run length
(beads spread from the replicated start), and the uneven-replica batch of 512
two-dimensional dipole systems in a trap (N in {8..128}, P in {16,32,64}, T in
[1,4] K) as replica-steps/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R]
                       [--plain-steps S] [--no-cpu-baseline] [--cpu-steps S]
The timed region is K batched steps (every replica advances K steps);
`value` is replica-steps/s, the plain form is reported beside it.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pimd-b_amd"))

import pimdb_amd as G  # noqa: E402

# include/units.h constants
KELVIN = 3.1668152e-06
ANGSTROM = 1.8897261
FS = 4.1341373e16 * 1e-15
AMU = 1822.8885

PEAK_FP64_TFLOPS = 78.6      # MI355X published FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def workload(N=256, P=64, T=1.5, density=0.02186, seed=12345, rmin_A=2.2):
    L = (N / density) ** (1.0 / 3.0) * ANGSTROM
    rng = np.random.default_rng(seed)
    rmin = rmin_A * ANGSTROM
    pts = []
    while len(pts) < N:
        c = rng.uniform(-0.5 * L, 0.5 * L, size=3)
        ok = True
        for q in pts:
            d = c - q
            d -= L * np.floor(d / L + 0.5)
            if d @ d < rmin * rmin:
                ok = False
                break
        if ok:
            pts.append(c)
    x0 = np.array(pts)
    x = np.repeat(x0[None], P, axis=0).copy()
    mass = 4.0026 * AMU
    temp = T * KELVIN
    dt = 0.5 * FS
    kw = dict(natoms=N, nbeads=P, ndim=3, mass=mass, temperature=temp, dt=dt, gamma=1.0 / (100.0 * dt),
              bosonic=1, pbc=1, size=L, ext_kind="free", int_kind="aziz", cutoff=0.5 * L,
              propagator="cartesian", thermostat="langevin", nmthermostat=1, fixcom=1, seed=4242)
    thermo_beta = 1.0 / temp / P
    p = mass * rng.normal(0.0, 1.0 / np.sqrt(thermo_beta * mass), size=x.shape)
    p -= p.mean(axis=(0, 1), keepdims=True)
    return kw, x, p


MEV = 0.036749326e-3


def uneven_dipole_replicas(n=512, seed=2024):
    """Uneven two-dimensional replicas: N drawn from {8, 16, ..., 128}, P from
    {16, 32, 64}, T uniform in [1, 4] K, He-4 mass, harmonic trap 3 meV plus
    dipole pairs with a strength drawn from U[0.5, 2] (atomic units), all
    pairs, per-replica seeds.  Positions: a jittered 4-bohr lattice (so the
    1/r^3 pairs stay integrable at dt = 20 a.u.) plus a 0.3-bohr bead spread;
    momenta N(0, 0.5).  Returns [(params kw, x, p)]."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        N = int(rng.choice(np.arange(8, 129, 8)))
        P = int(rng.choice([16, 32, 64]))
        T = rng.uniform(1.0, 4.0)
        kw = dict(natoms=N, nbeads=P, ndim=2, mass=4.0026 * AMU, temperature=T * KELVIN, dt=20.0,
                  gamma=1.0 / 2000.0, bosonic=1, pbc=0, size=60.0, ext_kind="harmonic", ext_omega=3.0 * MEV,
                  int_kind="dipole", int_strength=rng.uniform(0.5, 2.0), cutoff=-1.0, propagator="cartesian",
                  thermostat="langevin", nmthermostat=int(r % 2), fixcom=int(r % 3 == 0), seed=1000 + r)
        side = int(np.ceil(np.sqrt(N)))
        grid = np.array([(a, c) for a in range(side) for c in range(side)][:N], dtype=float) * 4.0
        centres = grid - grid.mean(axis=0) + rng.uniform(-0.5, 0.5, size=(N, 2))
        x = np.repeat(centres[None], P, axis=0) + rng.normal(0, 0.3, size=(P, N, 2))
        p = rng.normal(0, 0.5, size=(P, N, 2))
        out.append((kw, x, p))
    return out


def host_cpu():
    """CPU model and the cores this process may use."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, os.cpu_count(), avail


def pair_stats(x, L, cutoff):
    """Unordered pairs and pairs inside the cutoff (minimum image), all beads."""
    P, N, _ = x.shape
    n_in = 0
    iu = np.triu_indices(N, 1)
    for b in range(P):
        d = x[b][:, None, :] - x[b][None, :, :]
        d -= L * np.floor(d / L + 0.5)
        r = np.sqrt((d * d).sum(-1))[iu]
        n_in += int(np.count_nonzero(r < cutoff))
    return P * N * (N - 1) // 2, n_in


def flops_model(kw, n_pairs, n_in):
    """Algorithmic work of one MD step and of one force-kernel launch
    (counting: +,-,*,min,compare,floor 1; FMA 2; div, sqrt 1; min image
    5/axis; exp/log counted separately as transcendentals)."""
    N, P, D = kw["natoms"], kw["nbeads"], kw["ndim"]
    F_pairs = n_pairs * 8 * D + n_in * 48
    X_pairs = 2 * n_in
    F_springs = N * (P - 2) * D * 14
    F_exchange = N * N * (12 * D + 11)
    X_exchange = 1.5 * N * N
    F_integr = 15 * N * P * D
    F_nm = 8 * P * P * N * D
    # the batched kernel transforms only the noise, once per MD step (one
    # P x P product per column): 2 P^2 N D of the 8 P^2 N D counted
    F_nm_exec = 2 * P * P * N * D
    return dict(force_kernel_flops=F_pairs + F_springs, force_kernel_exps=X_pairs,
                step_flops=F_pairs + F_springs + F_exchange + F_integr + F_nm,
                step_flops_executed=F_pairs + F_springs + F_exchange + F_integr + F_nm_exec,
                nm_flops=F_nm, nm_flops_executed=F_nm_exec,
                step_exps=X_pairs + X_exchange, exchange_flops=F_exchange, exchange_exps=X_exchange)


def bench_plain(kw, x, p, steps, warmup):
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    ctx.forces()
    ctx.graph_capture()
    ctx.step(warmup)
    t0 = time.perf_counter()
    ctx.step(steps)          # returns after the stream is synchronised
    t1 = time.perf_counter()
    ms_forces, ms_exch, ms_step = ctx.time_kernels(reps=20)
    ctx.close()
    return (t1 - t0) / steps * 1e3, ms_forces, ms_exch, ms_step


def bench_batched(kw, x, p, replicas, steps, warmup, threads=0, resident=0, phases=True, run_length=0):
    """R copies of the workload (distinct noise seeds) in the persistent
    kernel.  Returns wall ms per batched step over the timed region, device ms
    per launch from stream events around the same region, the launch shape,
    finiteness, (optionally) per-phase device times and (optionally) the
    device ms per launch of `steps` steps timed after `run_length` more steps
    of evolution."""
    params = []
    for r in range(replicas):
        k = dict(kw)
        k["seed"] = kw["seed"] + 7919 * r
        params.append(G.make_params(**k))
    b = G.Batch(params)
    for r in range(replicas):
        b.set_state(r, x, p)
    b.forces()
    if threads or resident:
        b.configure(threads, resident)
        tuned = (threads, resident)
    else:
        tuned = b.autotune(trial_steps=2)
    b.step(warmup)
    t0 = time.perf_counter()
    dev_ms = b.step_timed(steps)     # returns after the closing event completed
    t1 = time.perf_counter()
    ph = None
    if phases:
        ph = dict(pairs=b.time_phase(1), exchange=b.time_phase(2), integrator=b.time_phase(3))
    later = None
    if run_length > 0:
        b.step(run_length)
        later = b.step_timed(steps) / steps
    xs, _, fs = b.state(0)
    ok = bool(np.all(np.isfinite(xs)) and np.all(np.isfinite(fs)))
    b.close()
    return (t1 - t0) / steps * 1e3, dev_ms / steps, tuned[0], tuned[1], ok, ph, later


def bench_uneven(steps, warmup):
    """The 512 uneven dipole replicas: replica-steps/s from stream events."""
    reps = uneven_dipole_replicas()
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    b.forces()
    threads, resident = b.autotune(trial_steps=2)
    b.step(warmup)
    dev_ms = b.step_timed(steps) / steps
    finite = all(np.all(np.isfinite(b.state(r)[0])) for r in (0, len(reps) - 1))
    b.close()
    work = dict(replicas=len(reps), natoms_sum=int(sum(kw["natoms"] for kw, _, _ in reps)),
                bead_particles=int(sum(kw["natoms"] * kw["nbeads"] for kw, _, _ in reps)))
    return dict(value=len(reps) * 1e3 / dev_ms, unit="replica-steps/s", ms_per_step=dev_ms, steps=steps,
                batch_threads=threads, batch_grid=min(len(reps), resident), batch_resident=resident,
                finite_state=bool(finite), workload="512 uneven replicas: N in {8..128}, P in {16,32,64}, "
                "T in U[1,4] K, D=2, He-4, trap 3 meV + dipole pairs (strength U[0.5,2] a.u.), all pairs, "
                "half with normal-mode Langevin, seeded (bench.uneven_dipole_replicas)", **work)


def traffic_from_profile():
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by the same
    counter's reading on a copy kernel of known size), or None."""
    path = os.path.join(REPO, "profiles", "k_batch_step_aziz_n256_p64_r512_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(kw, x, p, cpu_steps):
    """The oracle (single-thread C restatement of the reference path) on the
    host cores of this machine, on the same initial state; also the force
    parity of both GPU forms on that state."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    sim = O.Sim(O.make_params(**kw))
    sim.set_state(x, p)
    sim.forces()
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        sim.step()
    dt = time.perf_counter() - t0
    f_ref = O.update_forces(O.make_params(**kw), x)

    def err(f):
        return max(float(np.max(np.abs(f[b] - f_ref[b])) / np.max(np.abs(f_ref[b]))) for b in range(x.shape[0]))

    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    ctx.forces()
    e_plain = err(ctx.state()[2])
    ctx.close()
    bt = G.Batch([G.make_params(**kw)])
    bt.set_state(0, x, p)
    bt.forces()
    e_batch = err(bt.state(0)[2])
    bt.close()
    model, ncpu, avail = host_cpu()
    return dict(value=cpu_steps / dt, unit="MD steps/s", cores=1, kind="port",
                sample="%d OBABO steps of the benchmark workload from its initial state "
                       "(oracle/pimdb_oracle.c, one thread)" % cpu_steps,
                cpu_model=model, host_cpus=ncpu, host_cpus_available=avail), e_plain, e_batch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed batched MD steps")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--run-length", type=int, default=2000,
                    help="steps of evolution before the second timing (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-replica batch")
    ap.add_argument("--threads", type=int, default=0, help="batch workgroup size (0: autotune)")
    ap.add_argument("--resident", type=int, default=0, help="batch workgroups (0: autotune)")
    ap.add_argument("--no-phases", action="store_true", help="skip the per-phase timing launches")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the benchmark runs a ring polymer (or a set of "
              "replicas) on a single MI355X; --gpus %d is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    kw, x, p = workload()
    n_pairs, n_in = pair_stats(x, kw["size"], kw["cutoff"])
    fm = flops_model(kw, n_pairs, n_in)
    fma_tf, exp_g, copy_gbs = G.peak_probe()
    # the exp rate scaled to the nominal FP64 peak by the FMA rate measured
    # beside it, so that flops and exps are priced against the same machine
    exp_nominal = exp_g * (PEAK_FP64_TFLOPS / fma_tf)

    ms_plain, ms_forces, ms_exch, ms_step_graph = bench_plain(kw, x, p, args.plain_steps, 50)
    R = args.replicas
    ms_batch, ms_launch, threads, resident, finite, phases, ms_later = bench_batched(
        kw, x, p, R, args.steps, args.warmup, args.threads, args.resident, not args.no_phases, args.run_length)
    value = R * 1e3 / ms_batch
    grid = min(R, resident)

    def frac(flops, exps, ms):
        return (flops / (PEAK_FP64_TFLOPS * 1e12) + exps / (exp_nominal * 1e9)) / (ms * 1e-3)

    # dominant kernel: the persistent step kernel (one launch = one MD step of all R replicas)
    fl = R * fm["step_flops"]
    fl_exec = R * fm["step_flops_executed"]
    xp = R * fm["step_exps"]
    achieved = fl / (ms_launch * 1e-3) / 1e12
    roofline = dict(bound="valu_fp64", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP64_TFLOPS, traffic=traffic_from_profile(),
                    compute="FP64 on the vector ALUs: pairs, exchange, springs and the integrator's "
                            "elementwise work (%.0f%% of the counted flops); the normal-mode transform "
                            "(%.0f%% counted) on the matrix cores (v_mfma_f64_16x16x4f64), of which the "
                            "kernel executes a quarter (one noise product per step: C^T (c1 C p + c2 xi) = "
                            "c1 p + C^T c2 xi, two O-steps fused); MI355X's FP64 vector and matrix peaks "
                            "are both 78.6 TFLOP/s" % (100.0 * (1 - fm["nm_flops"] / fm["step_flops"]),
                                                       100.0 * fm["nm_flops"] / fm["step_flops"]),
                    kernel="k_batch_step<3,AZIZ,STEP> (persistent, %d replicas, grid %d x %d threads)"
                           % (R, grid, threads),
                    ms_per_launch=ms_launch, timing="hipEvents on the batch stream around the timed steps",
                    flops_per_launch=fl, flops_executed_per_launch=fl_exec,
                    frac_executed=fl_exec / (ms_launch * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                    exps_per_launch=xp, frac_with_exps=frac(fl, xp, ms_launch),
                    exp_pricing="exp at the measured rate scaled by 78.6 / measured FMA TFLOP/s "
                                "(%.0f Gexp/s)" % exp_nominal)
    achieved_f = fm["force_kernel_flops"] / (ms_forces * 1e-3) / 1e12

    batch_phases = None
    if phases:
        batch_phases = dict(
            ms=phases,
            pairs_frac=frac(R * fm["force_kernel_flops"], R * fm["force_kernel_exps"], phases["pairs"]),
            exchange_frac=frac(R * fm["exchange_flops"], R * fm["exchange_exps"], phases["exchange"]),
            exchange_frac_flops_only=R * fm["exchange_flops"] / (phases["exchange"] * 1e-3) / 1e12
            / PEAK_FP64_TFLOPS,
            exchange_tflops=R * fm["exchange_flops"] / (phases["exchange"] * 1e-3) / 1e12)

    out = dict(
        metric="MD steps/s N=256 P=64 Aziz; batched/persistent vs plain GPU @force err<=1e-10",
        value=value, unit="MD steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_batch, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload="N=256 He-4 bosons, P=64, D=3, T=1.5 K, Aziz, PBC 0.02186/A^3 "
                             "(L=42.914 bohr), cutoff L/2, dt=0.5 fs, NM Langevin, fixcom, FP64; positions "
                             "seeded uniform with a 2.2 A minimum-distance rejection, replicated over the "
                             "beads; every replica starts from the same x and p (Maxwell-Boltzmann) and "
                             "draws its own noise; timed from the start (warmup steps after it)",
                    form="batched persistent kernel + hipGraph, %d replicas (value = replica-steps/s)" % R,
                    replicas=R, natoms=256, nbeads=64, pairs_per_step=n_pairs, pairs_in_cutoff=n_in,
                    batch_threads=threads, batch_grid=grid, batch_resident=resident),
        roofline=roofline,
        per_replica_steps_per_s=1e3 / ms_batch,
        after_run_length=None if ms_later is None else dict(
            steps_before=args.warmup + args.steps + args.run_length, ms_per_step=ms_later,
            replica_steps_per_s=R * 1e3 / ms_later,
            note="the same batch timed again after the workload's run length (beads spread)"),
        plain=dict(steps_per_s=1e3 / ms_plain, ms_per_step=ms_plain, steps=args.plain_steps,
                   form="one system, per-kernel launches replayed from a hipGraph",
                   forces_kernel_ms=ms_forces, forces_kernel_tflops=achieved_f,
                   forces_kernel_frac=achieved_f / PEAK_FP64_TFLOPS, exchange_chain_ms=ms_exch,
                   step_graph_ms=ms_step_graph),
        batched_phases=batch_phases,
        measured_peaks=dict(fp64_fma_tflops=fma_tf, fp64_exp_gexp_s=exp_g, copy_gbs=copy_gbs),
        finite_state=finite,
    )
    if not args.no_uneven:
        out["uneven_replicas"] = bench_uneven(args.steps, args.warmup)
    if not args.no_cpu_baseline:
        cb, e_plain, e_batch = cpu_baseline(kw, x, p, args.cpu_steps)
        out["cpu_baseline"] = cb
        out["force_rel_err_vs_oracle"] = dict(plain=e_plain, batched=e_batch)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
