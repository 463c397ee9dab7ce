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
           kernel per MD step (when the library provides it)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R]
                       [--no-cpu-baseline] [--cpu-steps S]
"""
import argparse
import json
import os
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
    return dict(force_kernel_flops=F_pairs + F_springs, force_kernel_exps=X_pairs,
                step_flops=F_pairs + F_springs + F_exchange + F_integr + F_nm,
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


def cpu_baseline(kw, x, p, cpu_steps):
    """The oracle (single-thread C restatement of the reference path) on the
    host cores of this machine, on the same initial state."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    sim = O.Sim(O.make_params(**kw))
    sim.set_state(x, p)
    sim.forces()
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        sim.step()
    dt = time.perf_counter() - t0
    # force parity on the benchmark's own initial state
    f_ref = O.update_forces(O.make_params(**kw), x)
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x, p)
    ctx.forces()
    _, _, f_gpu = ctx.state()
    ctx.close()
    err = max(float(np.max(np.abs(f_gpu[b] - f_ref[b])) / np.max(np.abs(f_ref[b]))) for b in range(x.shape[0]))
    return dict(value=cpu_steps / dt, unit="MD steps/s", cores=1, kind="port",
                sample="%d OBABO steps of the benchmark workload from its initial state "
                       "(oracle/pimdb_oracle.c, one thread)" % cpu_steps), err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--replicas", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs a ring polymer (or a set of "
              "replicas) on a single MI355X; --gpus %d is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    kw, x, p = workload()
    n_pairs, n_in = pair_stats(x, kw["size"], kw["cutoff"])
    fm = flops_model(kw, n_pairs, n_in)
    fma_tf, exp_g, copy_gbs = G.peak_probe()

    ms_step, ms_forces, ms_exch, ms_step_ev = bench_plain(kw, x, p, args.steps, args.warmup)
    value = 1e3 / ms_step

    achieved = fm["force_kernel_flops"] / (ms_forces * 1e-3) / 1e12
    t_roof_forces = fm["force_kernel_flops"] / (PEAK_FP64_TFLOPS * 1e12) + fm["force_kernel_exps"] / (exp_g * 1e9)
    roofline = dict(bound="fp64", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP64_TFLOPS, traffic=None,
                    kernel="k_bead_forces<3,AZIZ>", ms_per_launch=ms_forces,
                    flops_per_launch=fm["force_kernel_flops"], exps_per_launch=fm["force_kernel_exps"],
                    frac_with_exps=t_roof_forces / (ms_forces * 1e-3))

    out = dict(
        metric="MD steps/s N=256 P=64 Aziz; batched/persistent vs plain GPU @force err<=1e-10",
        value=value, unit="MD steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_step, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload="N=256 He-4 bosons, P=64, D=3, T=1.5 K, Aziz, PBC 0.02186/A^3 "
                             "(L=42.914 bohr), cutoff L/2, dt=0.5 fs, NM Langevin, fixcom, FP64",
                    form="plain (per-kernel, hipGraph replay)", natoms=256, nbeads=64,
                    pairs_per_step=n_pairs, pairs_in_cutoff=n_in),
        roofline=roofline,
        stages_ms=dict(forces_kernel=ms_forces, exchange_chain=ms_exch, step_eager_events=ms_step_ev),
        measured_peaks=dict(fp64_fma_tflops=fma_tf, fp64_exp_gexp_s=exp_g, copy_gbs=copy_gbs),
        step_roofline=dict(flops=fm["step_flops"], exps=fm["step_exps"],
                           frac=(fm["step_flops"] / (PEAK_FP64_TFLOPS * 1e12) + fm["step_exps"] / (exp_g * 1e9))
                           / (ms_step * 1e-3)),
    )
    if not args.no_cpu_baseline:
        cb, err = cpu_baseline(kw, x, p, args.cpu_steps)
        out["cpu_baseline"] = cb
        out["force_rel_err_vs_oracle"] = err
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
