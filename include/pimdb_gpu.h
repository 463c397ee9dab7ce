/*
 * pimdb_gpu.h -- C ABI of the MI355X (gfx950) implementation of the PIMD-B++
 * per-step hot path: bosonic exchange (Feldman-Hirshberg), pair and spring
 * forces, and the OBABO step with a Langevin thermostat.
 *
 * All state of a ring polymer (P beads x N particles x D axes, FP64) stays
 * resident in device memory.  Host arrays are caller-owned and copied; their
 * layout is [bead][particle][axis], i.e. a stack of the reference's dVec
 * (include/common.h:227-387).  Device arrays are bead-major SoA [bead][axis]
 * [particle].  Units are atomic units, IPI convention (include/common.h:178-190).
 *
 * Errors: every call returns PIMDB_OK or an error code; pimdb_last_error()
 * describes the last failure of a context.  PIMDB_E_NONFINITE mirrors the
 * reference's std::overflow_error from the exchange recursions
 * (src/bosonic_exchange/quadratic_bosonic_exchange.cpp:92-97, 119-124).
 * A context is used by one host thread at a time and owns one HIP stream.
 * No function falls back to a CPU path: a missing or unusable GPU is an
 * error (PIMDB_E_HIP).
 */
#ifndef PIMDB_GPU_H
#define PIMDB_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pimdb_status {
    PIMDB_OK = 0,
    PIMDB_E_NONFINITE = 1, /* exchange potential not finite (overflow_error) */
    PIMDB_E_ARG = 2,       /* invalid argument (invalid_argument)            */
    PIMDB_E_HIP = 3,       /* HIP runtime failure or no GPU                   */
    PIMDB_E_STATE = 4      /* call not valid in the current state             */
};

/* Potentials (the classes of src/potentials).  DOUBLE_WELL and COSINE are external
 * potentials only, as in the reference's allowed names (params.cpp:189-190). */
enum pimdb_potential {
    PIMDB_POT_FREE = 0, PIMDB_POT_HARMONIC = 1, PIMDB_POT_DIPOLE = 2, PIMDB_POT_AZIZ = 3,
    PIMDB_POT_DOUBLE_WELL = 4, PIMDB_POT_COSINE = 5
};
enum pimdb_propagator { PIMDB_PROP_CARTESIAN = 0, PIMDB_PROP_NORMAL_MODES = 1 };
/* Nose-Hoover chains (nose_hoover.cpp): one chain per bead, per particle of
 * each bead (np) or per degree of freedom of each bead (np_dim). */
enum pimdb_thermostat { PIMDB_THERMO_NONE = 0, PIMDB_THERMO_LANGEVIN = 1, PIMDB_THERMO_NOSE_HOOVER = 2,
                        PIMDB_THERMO_NOSE_HOOVER_NP = 3, PIMDB_THERMO_NOSE_HOOVER_NP_DIM = 4 };

/* One ring-polymer system.  Replaces the Params maps and the Simulation
 * fields read from them (src/params.cpp:14-247, src/simulation.cpp:21-93).
 * Quantities are already in atomic units (temperature in Hartree). */
typedef struct pimdb_params {
    int natoms, nbeads, ndim;  /* ndim 1..3                                  */
    double mass;               /* electron masses                             */
    double temperature;        /* Hartree (kB = 1)                            */
    double dt, gamma;          /* time step, Langevin friction                */
    int bosonic, pbc;          /* bosonic exchange on beads 1 and P; PBC       */
    double size;               /* cubic box length (bohr)                      */
    int ext_kind;              /* FREE, HARMONIC, DOUBLE_WELL or COSINE         */
    double ext_omega;          /* trap frequency (energy units, hbar = 1)      */
    int int_kind;              /* pimdb_potential                              */
    double int_omega, int_strength, cutoff; /* cutoff < 0: all pairs           */
    int propagator, thermostat, nmthermostat, fixcom;
    unsigned int seed;         /* device noise stream seed                     */
    /* external double well: strength (energy), location (length); external
     * cosine: amplitude (energy), phase (number), wavelength = size
     * (double_well.cpp:3-4, cosine.cpp:4-7, simulation.cpp:624-637) */
    double ext_strength, ext_location, ext_amplitude, ext_phase;
    /* bosonic exchange algorithm: PIMDB_EXCHANGE_QUADRATIC (Feldman-Hirshberg,
     * quadratic_bosonic_exchange.cpp) or PIMDB_EXCHANGE_FACTORIAL (all N!
     * permutations, factorial_bosonic_exchange.cpp, natoms <= 10); the
     * reference chooses between them at build time
     * (FACTORIAL_BOSONIC_ALGORITHM, simulation.cpp:694-698) */
    int exchange_algo;
    /* Nose-Hoover chain length (the [simulation] key nchains, default 4,
     * params.cpp:22-25); used by the Nose-Hoover thermostats, >= 2 */
    int nchains;
} pimdb_params;

enum pimdb_exchange_algo { PIMDB_EXCHANGE_QUADRATIC = 0, PIMDB_EXCHANGE_FACTORIAL = 1 };

typedef struct pimdb_ctx pimdb_ctx;

/* Creates a single-system context (the "plain" per-kernel form).
 * Replaces Simulation::Simulation (src/simulation.cpp:16-101) minus
 * initial-condition sampling, which stays host code. */
int pimdb_ctx_create(const pimdb_params* params, pimdb_ctx** out);
void pimdb_ctx_destroy(pimdb_ctx* ctx);
/* Last failure of ctx; with ctx == NULL, the last failed create call of the
 * calling thread (the reference's invalid_argument messages, params.cpp). */
const char* pimdb_last_error(const pimdb_ctx* ctx);

/* Host <-> device state, [bead][particle][axis].  Any pointer may be NULL.
 * Replaces direct access to Simulation::coord/momenta/forces of every rank
 * (include/simulation.h:59-60). */
int pimdb_set_state(pimdb_ctx* ctx, const double* x, const double* p);
int pimdb_get_state(pimdb_ctx* ctx, double* x, double* p, double* f);

/* Forces on all beads from the current positions: exchange + interior
 * springs + physical forces.  Replaces Simulation::updateForces on every rank
 * (src/simulation.cpp:353-374) including BosonicExchange::prepare and
 * exteriorSpringForce (bosonic_exchange_base.cpp:72-78). */
int pimdb_forces(pimdb_ctx* ctx);

/* Exchange forward pass from the current positions of beads 1 and P.
 * VB = V[N] (effectivePotential, quadratic_bosonic_exchange.cpp:130-132);
 * V, Vb: N+1 values (evaluateVBn / evaluateVBackwards, :73-128);
 * prob: N x N row-major [l][u] (evaluateConnectionProbabilities, :142-157);
 * E: N(N+1)/2 packed as the reference's E_kn (getEnk, :63-66).
 * Any output may be NULL. */
int pimdb_exchange_forward(pimdb_ctx* ctx, double* VB, double* V, double* Vb, double* prob, double* E);

/* Exchange adjoint: exterior spring forces on beads 1 and P, N x D each
 * (springForceFirstBead / springForceLastBead, :159-215), from the positions
 * of beads 1, 2, P-1 and P.  Runs the forward pass itself. */
int pimdb_exchange_adjoint(pimdb_ctx* ctx, double* f_first, double* f_last);

/* nsteps OBABO steps (Simulation::run loop body, src/simulation.cpp:246-259).
 * host_noise: NULL for on-device Philox4x32-10 noise, or the Gaussian deviates
 * the reference's per-bead RanMars streams would draw, laid out
 * [step][ostep(2)][bead][particle][axis] (langevin.cpp:18-26, in mode order
 * under nmthermostat, thermostat_coupling.cpp:34-42).  Forces must be valid
 * on entry (pimdb_forces, or zero as the reference has them before step 0). */
int pimdb_step(pimdb_ctx* ctx, int nsteps, const double* host_noise);

/* Observables summed over beads, atomic units (EnergyObservable,
 * ClassicalObservable, BosonicObservable; src/observables/{energy,classical,
 * bosonic}.cpp): out[11] = {kinetic, potential, ext_pot, int_pot, virial,
 * temperature, cl_kinetic, cl_spring, prob_dist, prob_all, nh_energy}.
 * ext_pot/int_pot are 0 unless both potentials are non-free, as in
 * energy.cpp:96-102; nh_energy (the Nose-Hoover thermostats'
 * getAdditionToH summed over beads, classical.cpp:24-26) is 0 for the other
 * thermostats. */
int pimdb_observables(pimdb_ctx* ctx, double* out);

/* Exchange estimators from the current positions of beads 1 and P:
 * primEstimator (already divided by P under the IPI convention),
 * getDistinctProbability and getLongestProbability
 * (quadratic_bosonic_exchange.cpp:222-279).  Any output may be NULL;
 * PIMDB_E_STATE for a distinguishable system. */
int pimdb_exchange_estimators(pimdb_ctx* ctx, double* prim, double* prob_distinct, double* prob_longest);

/* Captures one MD step (device noise) in a hipGraph; later pimdb_step calls
 * with host_noise == NULL replay it.  No reference counterpart (the
 * reference has no graph capture). */
int pimdb_graph_capture(pimdb_ctx* ctx);

/* Average device time (ms) of `reps` launches of the physical-force kernel
 * of one force evaluation (ms_forces), of the exchange forward pass and
 * exterior forces (ms_exchange) and of one MD step with device noise
 * (ms_step), measured with HIP events on the context stream; any output may
 * be NULL.  Side effects: the forces are recomputed from the current
 * positions after the force and exchange timings, so those leave the state
 * as it was; ms_step advances the state by 2 * reps device-noise steps.
 * No reference counterpart (the reference times only the run loop,
 * src/simulation.cpp:225-226, 282-289). */
int pimdb_time_kernels(pimdb_ctx* ctx, int reps, double* ms_forces, double* ms_exchange, double* ms_step);

/* Micro-benchmarks for the roofline: FP64 FMA rate (TFLOP/s), FP64 exp
 * throughput of the pair kernels' exp (Gexp/s) and HBM copy bandwidth
 * (GB/s, read + write, over copy_bytes; <= 0 means 1 GiB).  Any output may
 * be NULL.  No reference counterpart. */
int pimdb_peak_probe(double* fma_tflops, double* exp_gexp_per_s, double* copy_gbs, long long copy_bytes);

/* Compute units of the current device and the LDS bytes one workgroup may
 * use (the batched form plans its replicas' phases for that LDS, shared by
 * the workgroups of one CU).  Any output may be NULL.  No reference
 * counterpart. */
int pimdb_device_info(int* compute_units, long long* lds_bytes);

/* Evaluates the device elementary functions on n host values:
 * fn 0 = e^x (exp_minimax, the exchange's and integrator's e^x), fn 1 =
 * 1/sqrt(x) to 1 ulp (rsqrt_refined, the full-precision form of fn 6), fn 2 = ln x
 * for x in (0, 1] (log_unit), fn 3 / 4 = sin / cos of 2 pi x for x in [0, 1)
 * (sincos_2pi; these three make the Box-Muller noise), fn 5 = e^x for
 * x <= 0 (exp_pair, the Aziz force's exponentials), fn 6 = 1/sqrt(x)
 * (rsqrt_pair, the pair forces' 1/r), y = f(x).  A test
 * hook for their accuracy.  No reference counterpart (the reference calls the
 * C library's exp and sqrt, aziz.cpp:56-106, and draws its noise on the host,
 * random_mars.cpp:83-103). */
int pimdb_math_probe(int fn, const double* x, double* y, long long n);

/* ---------------------------------------------------------------------- */
/* Exchange on caller-provided device arrays (the autograd harness).       */
/* x1, xP: device N x D SoA [axis][particle].  Outputs are device arrays:  */
/* VB (1), V (N+1), Vb (N+1), E (N(N+1)/2, may be NULL: then held inside). */
/* grad: device [2][D][N] = -dV_B/dx1, -dV_B/dxP (the exchange forces      */
/* without the interior links).  stream: a hipStream_t or NULL.            */
/* No reference counterpart: the reference has no differentiable API.      */
typedef struct pimdb_xplan pimdb_xplan;
int pimdb_xplan_create(int natoms, int ndim, double beta_t, double spring_k, int pbc, double size,
                       pimdb_xplan** out);
void pimdb_xplan_destroy(pimdb_xplan* plan);
int pimdb_xplan_forward(pimdb_xplan* plan, const double* x1, const double* xP, double* VB, double* V,
                        double* Vb, void* stream);
int pimdb_xplan_vb_grad(pimdb_xplan* plan, const double* x1, const double* xP, double* grad, void* stream);

/* pimdb_xplan_create + pimdb_xplan_forward (VB only) + pimdb_xplan_vb_grad in
 * one call that allocates nothing: the same kernels, with their workspace
 * (V, Vb, the ring's link prefix sums, the overflow flag) in a fixed device
 * array of the library (64 KB per GPU, held until the process ends).  x1, xP,
 * VB (may be NULL) and grad (may be NULL: forward only) are device arrays as
 * above.  natoms <= 2048 (and within the recursion's LDS workspace), else
 * PIMDB_E_ARG; the message of any failure is pimdb_last_error(NULL).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other threads) race.  Synchronises `stream` before returning.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc). */
int pimdb_exchange_vb_grad_noalloc(int natoms, int ndim, double beta_t, double spring_k, int pbc, double size,
                                   const double* x1, const double* xP, double* VB, double* grad, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched replicas: many independent systems of uneven N, P and T advanced */
/* by one persistent kernel per MD step (device noise), replayed from a     */
/* hipGraph.  All replicas share ndim and the interaction kind; they take  */
/* the Feldman-Hirshberg exchange, the Langevin thermostat, Nose-Hoover     */
/* chains (either coupling) or none, and either propagator (normal          */
/* modes without exchange, as the reference); each replica equals one      */
/* independent reference run                                                */
/* (the reference has no batched form).  A replica draws the device noise  */
/* stream of its seed, so it follows pimdb_step of a single-system context  */
/* with the same seed and state up to summation order.                      */
typedef struct pimdb_batch pimdb_batch;
int pimdb_batch_create(const pimdb_params* reps, int nrep, pimdb_batch** out);
void pimdb_batch_destroy(pimdb_batch* batch);
/* Last failure of batch; with batch == NULL, the last failed create call. */
const char* pimdb_batch_last_error(const pimdb_batch* batch);
/* Per-replica state, host layout [bead][particle][axis].  With fixcom and a
 * normal-mode Langevin thermostat on the matrix-core path (nbeads <= 64), the
 * COM subtraction after a step's last O-step is left pending on the device
 * and applied by the next step; get_state returns p with it applied. */
int pimdb_batch_set_state(pimdb_batch* batch, int rep, const double* x, const double* p);
int pimdb_batch_get_state(pimdb_batch* batch, int rep, double* x, double* p, double* f);
/* Forces of every replica (pimdb_forces for each). */
int pimdb_batch_forces(pimdb_batch* batch);
/* nsteps MD steps of every replica (pimdb_step with device noise for each). */
int pimdb_batch_step(pimdb_batch* batch, int nsteps);
/* Workgroup size of the persistent kernel (256, 512 or 1024; 0 keeps the
 * current one) and number of workgroups (0 keeps the current one; default
 * one per compute unit).  More workgroups than compute units share each CU's
 * LDS, and every replica's phases are planned for that share (an exchange
 * workspace that does not fit goes to device memory, allocated here);
 * PIMDB_E_ARG (the previous shape kept) when some replica does not fit it. */
int pimdb_batch_configure(pimdb_batch* batch, int threads, int resident);
/* Times trial_steps steps for each launch shape (256, 512 or 1024 threads,
 * 1, 2 or 4 workgroups per compute unit, at most 1024 threads per CU) whose
 * LDS plan fits, keeps the fastest and restores the state it started from. */
int pimdb_batch_autotune(pimdb_batch* batch, int trial_steps, int* threads_out, int* resident_out);
/* nsteps MD steps of every replica, as pimdb_batch_step, with the device
 * time between events recorded on the batch's stream just before the first
 * and just after the last step returned in device_ms (no reference
 * counterpart: measurement). */
int pimdb_batch_step_timed(pimdb_batch* batch, int nsteps, double* device_ms);
/* Average device time of one batched MD step over reps graph replays
 * (advances the state by reps + 1 steps). */
int pimdb_batch_time_step(pimdb_batch* batch, int reps, double* ms_step);
/* Average device time of one launch doing only phase 0 (full step), 1 (pair,
 * trap and spring forces), 2 (exchange: cycle energies, recursions, exterior
 * forces) or 3 (O-steps, kicks and drift) for every replica; the state is
 * saved before and restored after. */
int pimdb_batch_time_phase(pimdb_batch* batch, int phase, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
