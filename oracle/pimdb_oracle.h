/*
 * pimdb_oracle.h -- CPU restatement of the PIMD-B++ per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in pimd-b_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  It is never linked into the
 * product library and never used as a fallback for it.
 *
 * Every function restates one reference function in the reference's own
 * arithmetic order (no FMA contraction: the file is built with
 * -ffp-contract=off), so that on the same inputs it reproduces the reference
 * bit for bit wherever libm agrees.  Citations are path:line into higj/pimd-b.
 *
 * Conventions
 *   - atomic units; IPI convention (exchange beta = beta/P, k = m (P/beta)^2)
 *   - host layout of a bead: N x D doubles, particle-major (a dVec);
 *     a full ring polymer is [bead][particle][axis].
 *   - E_kn is the packed triangle of the reference, E(m,k) at m(m+1)/2 - k.
 *
 * Pinning status of each function is listed in the header of pimdb_oracle.c.
 */
#ifndef PIMDB_ORACLE_H
#define PIMDB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_E_NONFINITE = 1, ORC_E_ARG = 2 };

enum { ORC_POT_FREE = 0, ORC_POT_HARMONIC = 1, ORC_POT_DIPOLE = 2, ORC_POT_AZIZ = 3,
       ORC_POT_DOUBLE_WELL = 4, ORC_POT_COSINE = 5 };
enum { ORC_PROP_CARTESIAN = 0, ORC_PROP_NORMAL_MODES = 1 };
enum { ORC_THERMO_NONE = 0, ORC_THERMO_LANGEVIN = 1, ORC_THERMO_NOSE_HOOVER = 2, ORC_THERMO_NOSE_HOOVER_NP = 3,
       ORC_THERMO_NOSE_HOOVER_NP_DIM = 4 };

/* One ring-polymer system, atomic units (mirrors Params/Simulation fields,
 * src/simulation.cpp:21-59, src/params.cpp:14-247). */
typedef struct orc_params {
    int natoms, nbeads, ndim;
    double mass;        /* electron masses */
    double temperature; /* Hartree (kB = 1) */
    double dt, gamma;
    int bosonic, pbc;
    double size;        /* box length L (bohr) */
    int ext_kind;       /* ORC_POT_FREE, _HARMONIC, _DOUBLE_WELL or _COSINE */
    double ext_omega;
    int int_kind;       /* ORC_POT_* */
    double int_omega, int_strength, cutoff;
    int propagator, thermostat, nmthermostat, fixcom;
    unsigned int seed;
    /* external double well (strength: energy, location: length) and cosine
     * (amplitude: energy, phase: number; the wavelength is the box size) */
    double ext_strength, ext_location, ext_amplitude, ext_phase;
    int exchange_algo;  /* ORC_EXCHANGE_QUADRATIC or _FACTORIAL (FACTORIAL_BOSONIC_ALGORITHM) */
    int nchains;        /* Nose-Hoover chain length (params.cpp:22, default 4) */
} orc_params;

enum { ORC_EXCHANGE_QUADRATIC = 0, ORC_EXCHANGE_FACTORIAL = 1 };
#define ORC_FACTORIAL_MAXN 10

/* ---- factorial exchange (src/bosonic_exchange/factorial_bosonic_exchange.cpp) ---- */
double orc_factorial_e_shift(int N, int D, double k, int pbc, double L,
                             const double* x_first, const double* x_last);
double orc_factorial_vb(int N, int D, double beta_t, double k, int pbc, double L,
                        const double* x_first, const double* x_last);
void orc_factorial_force_last(int N, int D, double beta_t, double k, int pbc, double L,
                              const double* x_first, const double* x_last, const double* x_prev,
                              double e_shift, double* f);
void orc_factorial_force_first(int N, int D, double beta_t, double k, int pbc, double L,
                               const double* x_first, const double* x_last, const double* x_next,
                               double e_shift, double* f);
double orc_factorial_prim_estimator(int N, int P, int D, double beta_t, double k, int pbc, double L,
                                    const double* x_first, const double* x_last, double e_shift);

/* ---- exchange (src/bosonic_exchange/quadratic_bosonic_exchange.cpp) ---- */
double orc_min_image(double dx, double L);
void orc_cycle_energies(int N, int D, const double* x_first, const double* x_last,
                        double k, int pbc, double L, double* E_kn);
int orc_v_forward(int N, double beta_t, const double* E_kn, double* V);
int orc_v_backward(int N, double beta_t, const double* E_kn, const double* V, double* Vb);
void orc_connection_probabilities(int N, double beta_t, const double* E_kn, const double* V,
                                  const double* Vb, double* prob);
void orc_spring_force_first_bead(int N, int D, const double* x, const double* x_prev,
                                 const double* x_next, const double* prob, double k,
                                 int pbc, double L, double* f);
void orc_spring_force_last_bead(int N, int D, const double* x, const double* x_prev,
                                const double* x_next, const double* prob, double k,
                                int pbc, double L, double* f);
double orc_prim_estimator(int N, int P, double beta_t, const double* E_kn, const double* V);
double orc_distinct_probability(int N, double beta_t, const double* E_kn, const double* V);
double orc_longest_probability(int N, double beta_t, const double* E_kn, const double* V);
/* Convenience: E, V, Vb, prob and both exterior forces from beads 1, 2, P-1, P.
 * Any output may be NULL.  Returns ORC_E_NONFINITE like the overflow_error. */
int orc_exchange(int N, int D, double beta_t, double k, int pbc, double L,
                 const double* x1, const double* x2, const double* xPm1, const double* xP,
                 double* E_kn, double* V, double* Vb, double* prob,
                 double* f_first, double* f_last);

/* orc_exchange's V, Vb, probabilities and exterior forces computed in
 * __float128 (pimdb_oracle_quad.c): the exact values, to about 1e-30, for the
 * given double inputs; outputs rounded to double, any may be NULL. */
int orc_exchange_quad(int N, int D, double beta_t, double k, int pbc, double L,
                      const double* x1, const double* x2, const double* xPm1, const double* xP,
                      double* V_out, double* Vb_out, double* prob_out, double* f_first, double* f_last);
/* Brute force: V_B summed over all 2^(N-1) compositions of 0..N-1 into
 * consecutive cycles, -1/beta_t ln sum_c prod_blocks exp(-beta_t E)/(e+1). */
double orc_vb_enumerate(int N, double beta_t, const double* E_kn);

/* ---- potentials (src/potentials/{harmonic,dipole,aziz}.cpp) ---- */
void orc_harmonic_grad(int D, double mass, double omega, const double* x, double* g);
double orc_harmonic_V(int D, double mass, double omega, const double* x);
void orc_double_well_grad(int D, double mass, double strength, double loc, const double* x, double* g);
double orc_double_well_V(int N, int D, double mass, double strength, double loc, const double* coord);
void orc_cosine_grad(int D, double amplitude, double wavelength, double phase, const double* x, double* g);
double orc_cosine_V(int N, int D, double amplitude, double wavelength, double phase, const double* coord);
/* External potential of one bead (N x D) and its gradient at one particle. */
double orc_ext_V(const orc_params* p, const double* coord);
void orc_ext_grad(const orc_params* p, const double* x, double* g);
void orc_dipole_grad(int D, double strength, const double* x, double* g);
double orc_dipole_V(int D, double strength, const double* x);
void orc_aziz_grad(int D, const double* x, double* g);
double orc_aziz_V(int D, const double* x);

/* ---- forces (src/simulation.cpp:353-455) ---- */
double orc_effective_cutoff(const orc_params* p);
void orc_physical_forces(const orc_params* p, const double* coord, double* f);
void orc_interior_spring_forces(int N, int D, double k, int pbc, double L,
                                const double* coord, const double* prev,
                                const double* next, double* f);
/* Forces on every bead of a ring: f = spring + phys.  x, f are [P][N][D]. */
int orc_update_forces(const orc_params* p, const double* x, double* f);

/* ---- RNGs ---- */
typedef struct orc_ranmars {
    double u[98];
    int i97, j97, save;
    double c, cd, cm, second;
} orc_ranmars;
int orc_ranmars_init(orc_ranmars* r, int seed);
double orc_ranmars_uniform(orc_ranmars* r);
double orc_ranmars_gaussian(orc_ranmars* r);

typedef struct orc_mt19937 { unsigned int mt[624]; int mti; } orc_mt19937;
void orc_mt_seed(orc_mt19937* g, unsigned int seed);
unsigned int orc_mt_next(orc_mt19937* g);
double orc_mt_canonical(orc_mt19937* g);
double orc_mt_uniform_real(orc_mt19937* g, double a, double b);
double orc_mt_normal_fresh(orc_mt19937* g, double mean, double stddev);

/* ---- normal modes (src/normal_modes.cpp:47-81) ---- */
/* C[k][j]: cart->nm (row k as built on rank k); CT[b][k]: nm->cart (row b as
 * built on rank b).  Both P x P row-major. */
void orc_nm_matrices(int P, double* C, double* CT);

/* ---- whole-ring simulation state (src/simulation.cpp:16-101, 222-290) ---- */
typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(const orc_params* p);
void orc_sim_destroy(orc_sim* s);
/* random positions + Maxwell-Boltzmann momenta from mt19937(seed+bead),
 * then zeroMomentum (src/simulation.cpp:708-753). */
void orc_sim_init_random(orc_sim* s);
void orc_sim_set_state(orc_sim* s, const double* x, const double* p);
void orc_sim_get_state(const orc_sim* s, double* x, double* p, double* f);
/* Draws the noise of one MD step (2 O-steps) from the per-bead RanMars
 * streams, in consumption order: [ostep][bead][particle][axis]. */
void orc_sim_draw_noise(orc_sim* s, double* noise);
/* One OBABO MD step (src/simulation.cpp:246-259).  noise may be NULL (the
 * RanMars streams are then drawn in place).  Returns ORC_E_NONFINITE when the
 * exchange overflows. */
int orc_sim_step(orc_sim* s, const double* noise);
int orc_sim_forces(orc_sim* s);
void orc_sim_zero_momentum(orc_sim* s);
/* Observables summed over beads, atomic units:
 * [kinetic, potential, ext_pot, int_pot, virial, temperature(Hartree),
 *  cl_kinetic, cl_spring, prob_dist, prob_all, nh_energy]. */
int orc_sim_observables(orc_sim* s, double* out);

#ifdef __cplusplus
}
#endif
#endif
