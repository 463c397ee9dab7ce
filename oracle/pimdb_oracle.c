/*
 * pimdb_oracle.c -- CPU restatement of the PIMD-B++ per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pimdb_oracle.h).  Built with
 * -O2 -ffp-contract=off so every expression is evaluated as the reference's
 * x86-64 build evaluates it (baseline x86-64 has no FMA).
 *
 * Pinning of each part (tests/test_oracle_*.py):
 *   cycle energies, V, Vb, probabilities, exterior forces
 *       - reference fixtures x->F (bosonic_quadratic_harmonic_dynamics and
 *         _nmthermostat_dynamics: 101 + 1001 frames x 8 beads)
 *       - brute-force composition enumeration (orc_vb_enumerate) + central
 *         finite differences of V_B for the forces
 *       - closed forms N=1, N=2, coincident beads, distinguishable limit
 *       - row/column sums of the probability matrix
 *   potentials  - harmonic/dipole closed forms, Aziz gradient vs FD of V,
 *                 Aziz well depth V(r_m) = -epsilon; min image vs brute-force
 *                 image enumeration.  No reference fixture exercises pair
 *                 potentials, PBC or cutoffs: these are pinned only by
 *                 mathematics.
 *   springs, VV, Langevin, RanMars, mt19937 init, NM propagator/thermostat
 *       - full reference trajectories (positions, velocities, forces over
 *         1e5 steps) of the dynamics fixtures
 *       - RanMars vs the reference's own random_mars.cpp compiled in oracle/_ref
 *       - NM matrices orthogonality; free-ring exact rotation
 *   observables - simulation.out rows of the fixtures (rtol 1e-5)
 *   topology probabilities (orc_distinct_probability, orc_longest_probability)
 *       - their closed meaning over the compositions of the ring: the
 *         all-singleton share w(singletons)/Z and N w(one block)/Z, by
 *         enumeration for N <= 10 (tests/test_oracle_pins.py); N=1, N=2,
 *         coincident beads (1/N!, 1) and far-apart particles (1, 0)
 *   pair terms of the observables (int_pot, the r_i-only pair virial,
 *       energy.cpp:56-111) - worked by hand for 2-3 particles with harmonic,
 *       dipole and Aziz pairs, PBC and a pair past the cutoff
 *       (tests/test_oracle_pins.py)
 *   zeroMomentum - sums beads sequentially; the reference reduces them with
 *       MPI_Allreduce whose order is MPI-implementation defined.  All fixtures
 *       run with fixcom=false except the one initial call; that order is
 *       parity unpinned at the ulp level.
 */
#include "pimdb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* common.h:193 */
#define ORC_EPS 1.0E-7

/* src/common.cpp:41-43 */
double orc_min_image(double dx, double L) {
    dx -= L * floor(dx / L + 0.5);
    return dx;
}

/* BosonicExchangeBase::getBeadsSeparation, bosonic_exchange_base.cpp:30-42:
 * diff = x2[l2] - x1[l1], min-imaged. */
static void bead_sep(int N, int D, const double* x1, int l1, const double* x2, int l2,
                     int pbc, double L, double* diff) {
    l1 = l1 % N;
    l2 = l2 % N;
    for (int a = 0; a < D; ++a) {
        double dx = x2[l2 * D + a] - x1[l1 * D + a];
        if (pbc) dx = orc_min_image(dx, L);
        diff[a] = dx;
    }
}

/* bosonic_exchange_base.cpp:53-64 */
static double bead_sep_sq(int N, int D, const double* x1, int l1, const double* x2, int l2,
                          int pbc, double L) {
    double diff[8];
    bead_sep(N, D, x1, l1, x2, l2, pbc, L, diff);
    double d2 = 0.0;
    for (int a = 0; a < D; ++a) d2 += diff[a] * diff[a];
    return d2;
}

static inline double get_enk(const double* E, int m, int k) { return E[m * (m + 1) / 2 - k]; }
static inline void set_enk(double* E, int m, int k, double v) { E[m * (m + 1) / 2 - k] = v; }

/* quadratic_bosonic_exchange.cpp:34-61 (x_first = bead 1, x_last = bead P) */
void orc_cycle_energies(int N, int D, const double* x_first, const double* x_last,
                        double k, int pbc, double L, double* E_kn) {
    double* temp = (double*)malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; i++)
        temp[i] = bead_sep_sq(N, D, x_first, i, x_last, i, pbc, L);
    for (int v = 0; v < N; v++) {
        set_enk(E_kn, v + 1, 1, 0.5 * k * temp[v]);
        for (int u = v - 1; u >= 0; u--) {
            double val = get_enk(E_kn, v + 1, v - u) +
                0.5 * k * (
                    + bead_sep_sq(N, D, x_last, u, x_first, u + 1, pbc, L)
                    - bead_sep_sq(N, D, x_first, u + 1, x_last, v, pbc, L)
                    + bead_sep_sq(N, D, x_first, u, x_last, v, pbc, L));
            set_enk(E_kn, v + 1, v - u + 1, val);
        }
    }
    free(temp);
}

/* quadratic_bosonic_exchange.cpp:73-99 */
int orc_v_forward(int N, double beta, const double* E_kn, double* V) {
    double* temp = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    V[0] = 0.0;
    for (int m = 1; m < N + 1; m++) {
        double e_shift = DBL_MAX;
        for (int k = m; k > 0; k--) {
            double val = get_enk(E_kn, m, k) + V[m - k];
            if (val < e_shift) e_shift = val;   /* std::min(e_shift, val) */
            temp[k - 1] = val;
        }
        double sig_denom = 0.0;
        for (int k = m; k > 0; k--)
            sig_denom += exp(-beta * (temp[k - 1] - e_shift));
        V[m] = e_shift - (1.0 / beta) * log(sig_denom / (double)m);
        if (!isfinite(V[m])) { free(temp); return ORC_E_NONFINITE; }
    }
    free(temp);
    return ORC_OK;
}

/* quadratic_bosonic_exchange.cpp:101-128 */
int orc_v_backward(int N, double beta, const double* E_kn, const double* V, double* Vb) {
    double* temp = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    Vb[N] = 0.0;
    for (int l = N - 1; l > 0; l--) {
        double e_shift = DBL_MAX;
        for (int p = l; p < N; p++) {
            double val = get_enk(E_kn, p + 1, p - l + 1) + Vb[p + 1];
            if (val < e_shift) e_shift = val;
            temp[p] = val;
        }
        double sig_denom = 0.0;
        for (int p = l; p < N; p++)
            sig_denom += 1.0 / (p + 1) * exp(-beta * (temp[p] - e_shift));
        Vb[l] = e_shift - log(sig_denom) / beta;
        if (!isfinite(Vb[l])) { free(temp); return ORC_E_NONFINITE; }
    }
    Vb[0] = V[N];
    free(temp);
    return ORC_OK;
}

/* quadratic_bosonic_exchange.cpp:142-157.  prob is N x N, [l][u]; entries
 * never written by the reference (u > l+1) are set to 0 here, as the
 * reference's zero-initialised vector holds them. */
void orc_connection_probabilities(int N, double beta, const double* E_kn, const double* V,
                                  const double* Vb, double* prob) {
    memset(prob, 0, sizeof(double) * (size_t)N * (size_t)N);
    for (int l = 0; l < N - 1; l++) {
        double direct = 1.0 - (exp(-beta * (V[l + 1] + Vb[l + 1] - V[N])));
        prob[N * l + (l + 1)] = direct;
    }
    for (int u = 0; u < N; u++) {
        for (int l = u; l < N; l++) {
            double close = 1.0 / (l + 1) *
                exp(-beta * (V[u] + get_enk(E_kn, l + 1, l - u + 1) + Vb[l + 1] - V[N]));
            prob[N * l + u] = close;
        }
    }
}

/* quadratic_bosonic_exchange.cpp:188-215.  x = bead 1, x_prev = bead P,
 * x_next = bead 2 (all N x D). */
void orc_spring_force_first_bead(int N, int D, const double* x, const double* x_prev,
                                 const double* x_next, const double* prob, double k,
                                 int pbc, double L, double* f) {
    for (int l = 0; l < N; l++) {
        double sums[8] = {0};
        for (int prev_l = (l - 1 > 0 ? l - 1 : 0); prev_l < N; prev_l++) {
            double diff_prev[8];
            bead_sep(N, D, x, l, x_prev, prev_l, pbc, L, diff_prev);
            double pr = prob[N * prev_l + l];
            for (int a = 0; a < D; ++a) sums[a] += pr * diff_prev[a];
        }
        double diff_next[8];
        bead_sep(N, D, x, l, x_next, l, pbc, L, diff_next);
        for (int a = 0; a < D; ++a) sums[a] += diff_next[a];
        for (int a = 0; a < D; ++a) f[l * D + a] = sums[a] * k;
    }
}

/* quadratic_bosonic_exchange.cpp:159-186.  x = bead P, x_prev = bead P-1,
 * x_next = bead 1. */
void orc_spring_force_last_bead(int N, int D, const double* x, const double* x_prev,
                                const double* x_next, const double* prob, double k,
                                int pbc, double L, double* f) {
    for (int l = 0; l < N; l++) {
        double sums[8] = {0};
        for (int next_l = 0; next_l <= l + 1 && next_l < N; next_l++) {
            double diff_next[8];
            bead_sep(N, D, x, l, x_next, next_l, pbc, L, diff_next);
            double pr = prob[N * l + next_l];
            for (int a = 0; a < D; ++a) sums[a] += pr * diff_next[a];
        }
        double diff_prev[8];
        bead_sep(N, D, x, l, x_prev, l, pbc, L, diff_prev);
        for (int a = 0; a < D; ++a) sums[a] += diff_prev[a];
        for (int a = 0; a < D; ++a) f[l * D + a] = sums[a] * k;
    }
}

/* quadratic_bosonic_exchange.cpp:250-279 (IPI convention: / nbeads) */
double orc_prim_estimator(int N, int P, double beta, const double* E_kn, const double* V) {
    double* prim = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    prim[0] = 0.0;
    for (int m = 1; m < N + 1; ++m) {
        double sig = 0.0;
        double e_shift = DBL_MAX;
        for (int k = m; k > 0; k--) {
            double t = get_enk(E_kn, m, k) + V[m - k];
            if (t < e_shift) e_shift = t;
        }
        for (int k = m; k > 0; --k) {
            const double e = get_enk(E_kn, m, k);
            sig += (prim[m - k] - e) * exp(-beta * (e + V[m - k] - e_shift));
        }
        const double sig_denom_m = m * exp(-beta * (V[m] - e_shift));
        prim[m] = sig / sig_denom_m;
    }
    double r = prim[N] / P;
    free(prim);
    return r;
}

/* quadratic_bosonic_exchange.cpp:222-229 */
double orc_distinct_probability(int N, double beta, const double* E_kn, const double* V) {
    double s = 0.0;
    for (int m = 1; m < N + 1; ++m) s += get_enk(E_kn, m, 1);
    return exp(-beta * (s - V[N]) - lgamma(N + 1));
}

/* quadratic_bosonic_exchange.cpp:238-240 */
double orc_longest_probability(int N, double beta, const double* E_kn, const double* V) {
    return exp(-beta * (get_enk(E_kn, N, N) - V[N]));
}

int orc_exchange(int N, int D, double beta_t, double k, int pbc, double L,
                 const double* x1, const double* x2, const double* xPm1, const double* xP,
                 double* E_out, double* V_out, double* Vb_out, double* prob_out,
                 double* f_first, double* f_last) {
    if (N < 1) return ORC_E_ARG;
    size_t ne = (size_t)N * (N + 1) / 2;
    double* E = (double*)malloc(sizeof(double) * ne);
    double* V = (double*)malloc(sizeof(double) * (N + 1));
    double* Vb = (double*)malloc(sizeof(double) * (N + 1));
    double* pr = (double*)malloc(sizeof(double) * (size_t)N * N);
    int rc = ORC_OK;
    orc_cycle_energies(N, D, x1, xP, k, pbc, L, E);
    rc = orc_v_forward(N, beta_t, E, V);
    if (rc == ORC_OK) rc = orc_v_backward(N, beta_t, E, V, Vb);
    if (rc == ORC_OK) {
        orc_connection_probabilities(N, beta_t, E, V, Vb, pr);
        if (f_first) orc_spring_force_first_bead(N, D, x1, xP, x2, pr, k, pbc, L, f_first);
        if (f_last) orc_spring_force_last_bead(N, D, xP, xPm1, x1, pr, k, pbc, L, f_last);
        if (E_out) memcpy(E_out, E, sizeof(double) * ne);
        if (V_out) memcpy(V_out, V, sizeof(double) * (N + 1));
        if (Vb_out) memcpy(Vb_out, Vb, sizeof(double) * (N + 1));
        if (prob_out) memcpy(prob_out, pr, sizeof(double) * (size_t)N * N);
    }
    free(E); free(V); free(Vb); free(pr);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Factorial exchange, src/bosonic_exchange/factorial_bosonic_exchange.cpp:  */
/* the compile-time alternative over all N! permutations.  `labels` runs     */
/* through the permutations in lexicographic order from the identity          */
/* (std::ranges::next_permutation); the last bead of particle l connects to   */
/* the first bead of particle labels[l] (lastBeadNeighbor, :48-50), and the   */
/* first bead of l to the last bead of the particle at l's position in        */
/* labels (firstBeadNeighbor, :33-35).                                        */

/* std::next_permutation on a[0..n): the next permutation in lexicographic
 * order; after the last one, a is sorted again and 0 is returned. */
static int next_perm(int n, int* a) {
    int i = n - 2;
    while (i >= 0 && a[i] >= a[i + 1]) --i;
    if (i < 0) {
        for (int lo = 0, hi = n - 1; lo < hi; ++lo, --hi) { int t = a[lo]; a[lo] = a[hi]; a[hi] = t; }
        return 0;
    }
    int j = n - 1;
    while (a[j] <= a[i]) --j;
    int t = a[i]; a[i] = a[j]; a[j] = t;
    for (int lo = i + 1, hi = n - 1; lo < hi; ++lo, --hi) { t = a[lo]; a[lo] = a[hi]; a[hi] = t; }
    return 1;
}

static int first_bead_neighbor(int n, const int* labels, int l) {
    for (int m = 0; m < n; ++m)
        if (labels[m] == l) return m;
    return n;
}

/* getMinExteriorSpringEnergy, :57-79 */
double orc_factorial_e_shift(int N, int D, double k, int pbc, double L,
                             const double* x_first, const double* x_last) {
    int labels[ORC_FACTORIAL_MAXN];
    for (int i = 0; i < N; ++i) labels[i] = i;
    double min_delta = DBL_MAX;
    do {
        double diff2 = 0.0;
        for (int l = 0; l < N; ++l) diff2 += bead_sep_sq(N, D, x_first, l, x_last, labels[l], pbc, L);
        min_delta = fmin(min_delta, 0.5 * k * diff2);
    } while (next_perm(N, labels));
    return min_delta;
}

/* effectivePotential, :88-114 (beta_half_k = thermo_beta * 0.5 * k,
 * simulation.cpp:52) */
double orc_factorial_vb(int N, int D, double beta_t, double k, int pbc, double L,
                        const double* x_first, const double* x_last) {
    int labels[ORC_FACTORIAL_MAXN];
    for (int i = 0; i < N; ++i) labels[i] = i;
    const double beta_half_k = beta_t * 0.5 * k;
    long count = 0;
    double weights_sum = 0.0;
    do {
        count++;
        double diff2 = 0.0;
        for (int l = 0; l < N; ++l) diff2 += bead_sep_sq(N, D, x_first, l, x_last, labels[l], pbc, L);
        weights_sum += exp(-beta_half_k * diff2);
    } while (next_perm(N, labels));
    return (-1.0 / beta_t) * log(weights_sum / count);
}

/* springForceLastBead, :121-172: x_last = this bead (P), x_prev = bead P-1,
 * x_first = bead 1 (x_next of bead P). */
void orc_factorial_force_last(int N, int D, double beta_t, double k, int pbc, double L,
                              const double* x_first, const double* x_last, const double* x_prev,
                              double e_shift, double* f) {
    int labels[ORC_FACTORIAL_MAXN];
    double temp[ORC_FACTORIAL_MAXN * 8];
    for (int i = 0; i < N; ++i) labels[i] = i;
    for (int e = 0; e < N * D; ++e) f[e] = 0.0;
    double denom = 0.0;
    do {
        double weight = 0.0;
        for (int l = 0; l < N; ++l) {
            double dn[8], dp[8];
            bead_sep(N, D, x_last, l, x_first, labels[l], pbc, L, dn);
            for (int a = 0; a < D; ++a) weight += dn[a] * dn[a];
            bead_sep(N, D, x_last, l, x_prev, l, pbc, L, dp);
            for (int a = 0; a < D; ++a) temp[l * D + a] = (0.0 + dp[a] + dn[a]) * k;
        }
        weight = exp(-beta_t * (0.5 * k * weight - e_shift));
        for (int e = 0; e < N * D; ++e) f[e] += weight * temp[e];
        denom += weight;
    } while (next_perm(N, labels));
    for (int e = 0; e < N * D; ++e) f[e] = f[e] / denom;
}

/* springForceFirstBead, :179-230: x_first = this bead (1), x_last = bead P
 * (x_prev of bead 1), x_next = bead 2. */
void orc_factorial_force_first(int N, int D, double beta_t, double k, int pbc, double L,
                               const double* x_first, const double* x_last, const double* x_next,
                               double e_shift, double* f) {
    int labels[ORC_FACTORIAL_MAXN];
    double temp[ORC_FACTORIAL_MAXN * 8];
    for (int i = 0; i < N; ++i) labels[i] = i;
    for (int e = 0; e < N * D; ++e) f[e] = 0.0;
    double denom = 0.0;
    do {
        double weight = 0.0;
        for (int l = 0; l < N; ++l) {
            double dp[8], dn[8];
            bead_sep(N, D, x_first, l, x_last, first_bead_neighbor(N, labels, l), pbc, L, dp);
            for (int a = 0; a < D; ++a) weight += dp[a] * dp[a];
            bead_sep(N, D, x_first, l, x_next, l, pbc, L, dn);
            for (int a = 0; a < D; ++a) temp[l * D + a] = (0.0 + dp[a] + dn[a]) * k;
        }
        weight = exp(-beta_t * (0.5 * k * weight - e_shift));
        for (int e = 0; e < N * D; ++e) f[e] += weight * temp[e];
        denom += weight;
    } while (next_perm(N, labels));
    for (int e = 0; e < N * D; ++e) f[e] = f[e] / denom;
}

/* primEstimator, :258-290 (IPI convention: divided by P) */
double orc_factorial_prim_estimator(int N, int P, int D, double beta_t, double k, int pbc, double L,
                                    const double* x_first, const double* x_last, double e_shift) {
    int labels[ORC_FACTORIAL_MAXN];
    for (int i = 0; i < N; ++i) labels[i] = i;
    double numerator = 0.0, denom = 0.0;
    do {
        double weight = 0.0;
        for (int l = 0; l < N; ++l)
            weight += bead_sep_sq(N, D, x_first, l, x_last, first_bead_neighbor(N, labels, l), pbc, L);
        const double delta = 0.5 * k * weight;
        weight = exp(-beta_t * (delta - e_shift));
        numerator += delta * weight;
        denom += weight;
    } while (next_perm(N, labels));
    double e = numerator / denom;
    e /= P;
    return (-1.0) * e;
}

/* Brute force over compositions (independent of the recursion above): every
 * bit pattern of N-1 "cuts" splits 0..N-1 into consecutive blocks [s..e]; a
 * block weighs exp(-beta E^(s,e)) / (e+1) with E^(s,e) = E(e+1, e-s+1). */
double orc_vb_enumerate(int N, double beta, const double* E_kn) {
    if (N < 1 || N > 24) return NAN;
    /* shift by the minimal-energy composition to stay in range */
    unsigned long ncomp = 1UL << (N - 1);
    double emin = DBL_MAX;
    for (unsigned long c = 0; c < ncomp; ++c) {
        double e = 0.0; int s = 0;
        for (int i = 0; i < N; ++i) {
            int cut = (i == N - 1) || ((c >> i) & 1UL);
            if (cut) { e += get_enk(E_kn, i + 1, i - s + 1); s = i + 1; }
        }
        if (e < emin) emin = e;
    }
    double sum = 0.0;
    for (unsigned long c = 0; c < ncomp; ++c) {
        double e = 0.0, w = 1.0; int s = 0;
        for (int i = 0; i < N; ++i) {
            int cut = (i == N - 1) || ((c >> i) & 1UL);
            if (cut) { e += get_enk(E_kn, i + 1, i - s + 1); w /= (double)(i + 1); s = i + 1; }
        }
        sum += w * exp(-beta * (e - emin));
    }
    return emin - log(sum) / beta;
}

/* ------------------------------------------------------------------------ */
/* Potentials                                                               */

static double vnorm(int D, const double* x) {   /* VectorArray::norm, common.h:269 */
    double s = 0.0;
    for (int a = 0; a < D; a++) s += x[a] * x[a];
    return sqrt(s);
}

/* double_well.cpp:6-40: V sums (x_a^2 - loc^2)^2 over particles and axes,
 * times m*strength; gradV is the isotropic form 4 m s (|x|^2 - loc^2) x. */
void orc_double_well_grad(int D, double mass, double strength, double loc, const double* x, double* g) {
    const double loc2 = loc * loc;
    double pref = 0.0;
    for (int a = 0; a < D; ++a) pref += x[a] * x[a];
    pref = 4 * mass * strength * (pref - loc2);
    for (int a = 0; a < D; ++a) g[a] = pref * x[a];
}
double orc_double_well_V(int N, int D, double mass, double strength, double loc, const double* coord) {
    const double loc2 = loc * loc;
    double pot = 0.0;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < D; ++a) {
            const double x = coord[i * D + a];
            pot += (x * x - loc2) * (x * x - loc2);
        }
    pot *= mass * strength;
    return pot;
}

/* cosine.cpp:4-36: k = 2 pi / wavelength (the box size, simulation.cpp:634-637);
 * V = A sum cos(k x + phase); gradV = -A k sin(k x + phase). */
static const double ORC_PI = 3.14159265358979323846;
void orc_cosine_grad(int D, double amplitude, double wavelength, double phase, const double* x, double* g) {
    const double k = 2 * ORC_PI / wavelength;
    const double pref = -amplitude * k;
    for (int a = 0; a < D; ++a) g[a] = pref * sin(k * x[a] + phase);
}
double orc_cosine_V(int N, int D, double amplitude, double wavelength, double phase, const double* coord) {
    const double k = 2 * ORC_PI / wavelength;
    double pot = 0.0;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < D; ++a) pot += cos(k * coord[i * D + a] + phase);
    pot *= amplitude;
    return pot;
}

/* harmonic.cpp:3-27: k = m w^2, gradV = k x */
void orc_harmonic_grad(int D, double mass, double omega, const double* x, double* g) {
    double k = mass * omega * omega;
    for (int a = 0; a < D; ++a) g[a] = x[a] * k;
}
double orc_harmonic_V(int D, double mass, double omega, const double* x) {
    double k = mass * omega * omega, pot = 0.0;
    for (int a = 0; a < D; ++a) pot += x[a] * x[a];
    pot *= 0.5 * k;
    return pot;
}

/* dipole.cpp:24-45 */
void orc_dipole_grad(int D, double strength, const double* x, double* g) {
    const double norm = vnorm(D, x);
    const double pref = -3.0 * strength / (norm * norm * norm * norm * norm);
    for (int a = 0; a < D; ++a) g[a] = pref * x[a];
}
double orc_dipole_V(int D, double strength, const double* x) {
    const double norm = vnorm(D, x);
    return strength / (norm * norm * norm);
}

/* aziz.cpp:3-16 (HFDHE2) */
static const double AZ_RM = 5.60738, AZ_A = 0.5448504e6, AZ_EPS = 3.42016E-5,
    AZ_ALPHA = 13.353384, AZ_D = 1.241314, AZ_C6 = 1.3732412, AZ_C8 = 0.4253785,
    AZ_C10 = 0.1781;

/* aziz.h:25-43 */
static double aziz_F(double x) {
    return (x < AZ_D ? exp(-(AZ_D / x - 1.0) * (AZ_D / x - 1.0)) : 1.0);
}
static double aziz_dF(double x) {
    double ix = 1.0 / x;
    double r = 2.0 * AZ_D * ix * ix * (AZ_D * ix - 1.0) * exp(-(AZ_D * ix - 1.0) * (AZ_D * ix - 1.0));
    return (x < AZ_D ? r : 0.0);
}

/* aziz.cpp:56-106 */
void orc_aziz_grad(int D, const double* x, double* g) {
    const double norm = vnorm(D, x);
    const double r_scaled = norm / AZ_RM;
    const double T1 = -AZ_A * AZ_ALPHA * exp(-AZ_ALPHA * r_scaled);
    double grad_result;
    if (r_scaled > ORC_EPS && r_scaled < 0.01) {
        grad_result = T1 * (AZ_EPS / AZ_RM);
    } else {
        const double ix = 1.0 / r_scaled;
        const double ix2 = ix * ix;
        const double ix6 = ix2 * ix2 * ix2;
        const double ix7 = ix6 * ix;
        const double ix8 = ix6 * ix2;
        const double ix9 = ix8 * ix;
        const double ix10 = ix8 * ix2;
        const double ix11 = ix10 * ix;
        const double T2 = (6.0 * AZ_C6 * ix7 + 8.0 * AZ_C8 * ix9 + 10.0 * AZ_C10 * ix11) * aziz_F(r_scaled);
        const double T3 = -(AZ_C6 * ix6 + AZ_C8 * ix8 + AZ_C10 * ix10) * aziz_dF(r_scaled);
        grad_result = (AZ_EPS / AZ_RM) * (T1 + T2 + T3);
    }
    grad_result /= norm;
    for (int a = 0; a < D; ++a) g[a] = grad_result * x[a];
}

/* aziz.cpp:18-54 */
double orc_aziz_V(int D, const double* x) {
    const double r_scaled = vnorm(D, x) / AZ_RM;
    const double Urep = AZ_A * exp(-AZ_ALPHA * r_scaled);
    if (r_scaled > ORC_EPS && r_scaled < 0.01) return AZ_EPS * Urep;
    const double ix2 = 1.0 / (r_scaled * r_scaled);
    const double ix6 = ix2 * ix2 * ix2;
    const double ix8 = ix6 * ix2;
    const double ix10 = ix8 * ix2;
    const double Uatt = -(AZ_C6 * ix6 + AZ_C8 * ix8 + AZ_C10 * ix10) * aziz_F(r_scaled);
    return AZ_EPS * (Urep + Uatt);
}

static void int_grad(const orc_params* p, const double* diff, double* g) {
    switch (p->int_kind) {
    case ORC_POT_HARMONIC: orc_harmonic_grad(p->ndim, p->mass, p->int_omega, diff, g); break;
    case ORC_POT_DIPOLE: orc_dipole_grad(p->ndim, p->int_strength, diff, g); break;
    case ORC_POT_AZIZ: orc_aziz_grad(p->ndim, diff, g); break;
    default: for (int a = 0; a < p->ndim; ++a) g[a] = 0.0; break;
    }
}
static double int_V(const orc_params* p, const double* diff) {
    switch (p->int_kind) {
    case ORC_POT_HARMONIC: return orc_harmonic_V(p->ndim, p->mass, p->int_omega, diff);
    case ORC_POT_DIPOLE: return orc_dipole_V(p->ndim, p->int_strength, diff);
    case ORC_POT_AZIZ: return orc_aziz_V(p->ndim, diff);
    default: return 0.0;
    }
}

/* simulation.cpp:83-90 */
double orc_effective_cutoff(const orc_params* p) {
    double c = (p->int_kind == ORC_POT_FREE) ? 0.0 : p->cutoff;
    if (p->pbc) c = (c < 0.5 * p->size) ? c : 0.5 * p->size;   /* std::min */
    return c;
}

/* Simulation::initializePotential (simulation.cpp:612-645), external side */
void orc_ext_grad(const orc_params* p, const double* x, double* g) {
    const int D = p->ndim;
    switch (p->ext_kind) {
    case ORC_POT_HARMONIC: orc_harmonic_grad(D, p->mass, p->ext_omega, x, g); break;
    case ORC_POT_DOUBLE_WELL: orc_double_well_grad(D, p->mass, p->ext_strength, p->ext_location, x, g); break;
    case ORC_POT_COSINE: orc_cosine_grad(D, p->ext_amplitude, p->size, p->ext_phase, x, g); break;
    default: for (int a = 0; a < D; ++a) g[a] = 0.0; break;
    }
}
double orc_ext_V(const orc_params* p, const double* coord) {
    const int N = p->natoms, D = p->ndim;
    switch (p->ext_kind) {
    case ORC_POT_HARMONIC: {
        /* HarmonicPotential::V over the whole dVec */
        double vv = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) vv += coord[i * D + a] * coord[i * D + a];
        vv *= 0.5 * (p->mass * p->ext_omega * p->ext_omega);
        return vv;
    }
    case ORC_POT_DOUBLE_WELL: return orc_double_well_V(N, D, p->mass, p->ext_strength, p->ext_location, coord);
    case ORC_POT_COSINE: return orc_cosine_V(N, D, p->ext_amplitude, p->size, p->ext_phase, coord);
    default: return 0.0;
    }
}

/* simulation.cpp:499-512 (r1 - r2) */
static void pair_sep(const orc_params* p, const double* coord, int i, int j, double* d) {
    for (int a = 0; a < p->ndim; ++a) {
        double diff = coord[i * p->ndim + a] - coord[j * p->ndim + a];
        if (p->pbc) diff = orc_min_image(diff, p->size);
        d[a] = diff;
    }
}

/* simulation.cpp:428-455 */
void orc_physical_forces(const orc_params* p, const double* coord, double* f) {
    const int N = p->natoms, D = p->ndim;
    double g[8];
    for (int i = 0; i < N; ++i) {
        orc_ext_grad(p, coord + i * D, g);
        for (int a = 0; a < D; ++a) f[i * D + a] = g[a] * (-1.0);
    }
    const double cutoff = orc_effective_cutoff(p);
    if (cutoff != 0.0) {
        for (int i = 0; i < N; ++i) {
            for (int j = i + 1; j < N; ++j) {
                double diff[8];
                pair_sep(p, coord, i, j, diff);
                const double distance = vnorm(D, diff);
                if (distance < cutoff || cutoff < 0.0) {
                    int_grad(p, diff, g);
                    for (int a = 0; a < D; ++a) {
                        double fo = g[a] * (-1.0);
                        f[i * D + a] += fo;
                        f[j * D + a] -= fo;
                    }
                }
            }
        }
    }
}

/* simulation.cpp:405-419 */
void orc_interior_spring_forces(int N, int D, double k, int pbc, double L,
                                const double* coord, const double* prev,
                                const double* next, double* f) {
    for (int i = 0; i < N; ++i) {
        for (int a = 0; a < D; ++a) {
            double dp = prev[i * D + a] - coord[i * D + a];
            double dn = next[i * D + a] - coord[i * D + a];
            if (pbc) { dp = orc_min_image(dp, L); dn = orc_min_image(dn, L); }
            f[i * D + a] = k * (dp + dn);
        }
    }
}

typedef struct derived {
    double beta, thermo_beta, omega_p, k, beta_t;
} derived;

/* simulation.cpp:37-52; bosonic_exchange_base.cpp:16-18 */
static derived derive(const orc_params* p) {
    derived d;
    d.beta = 1.0 / (1.0 * p->temperature);
    d.thermo_beta = d.beta / p->nbeads;
    d.omega_p = p->nbeads / (d.beta * 1.0);
    d.k = p->mass * d.omega_p * d.omega_p;
    d.beta_t = d.beta / p->nbeads;
    return d;
}

static int is_bosonic(const orc_params* p) { return p->bosonic && p->nbeads > 1; }

/* simulation.cpp:353-420 for every bead of the ring */
int orc_update_forces(const orc_params* p, const double* x, double* f) {
    const int N = p->natoms, D = p->ndim, P = p->nbeads;
    const size_t nb = (size_t)N * D;
    derived d = derive(p);
    double* spring = (double*)malloc(sizeof(double) * nb);
    double* phys = (double*)malloc(sizeof(double) * nb);
    double* f_first = NULL; double* f_last = NULL;
    int rc = ORC_OK;
    if (is_bosonic(p)) {
        f_first = (double*)malloc(sizeof(double) * nb);
        f_last = (double*)malloc(sizeof(double) * nb);
        if (p->exchange_algo == ORC_EXCHANGE_FACTORIAL) {
            if (N > ORC_FACTORIAL_MAXN) {
                rc = ORC_E_ARG;
            } else {
                /* prepare() then exteriorSpringForce, simulation.cpp:396-401 */
                const double* x1 = x;
                const double* xP = x + nb * (P - 1);
                const double e_shift = orc_factorial_e_shift(N, D, d.k, p->pbc, p->size, x1, xP);
                orc_factorial_force_first(N, D, d.beta_t, d.k, p->pbc, p->size, x1, xP, x + nb * 1, e_shift,
                                          f_first);
                orc_factorial_force_last(N, D, d.beta_t, d.k, p->pbc, p->size, x1, xP, x + nb * (P - 2), e_shift,
                                         f_last);
            }
        } else {
            rc = orc_exchange(N, D, d.beta_t, d.k, p->pbc, p->size,
                              x, x + nb * 1, x + nb * (P - 2), x + nb * (P - 1),
                              NULL, NULL, NULL, NULL, f_first, f_last);
        }
    }
    if (rc == ORC_OK) {
        for (int b = 0; b < P; ++b) {
            const double* coord = x + nb * b;
            const double* prev = x + nb * ((b - 1 + P) % P);
            const double* next = x + nb * ((b + 1) % P);
            if (is_bosonic(p) && b == 0) memcpy(spring, f_first, sizeof(double) * nb);
            else if (is_bosonic(p) && b == P - 1) memcpy(spring, f_last, sizeof(double) * nb);
            else orc_interior_spring_forces(N, D, d.k, p->pbc, p->size, coord, prev, next, spring);
            orc_physical_forces(p, coord, phys);
            for (size_t e = 0; e < nb; ++e) f[nb * b + e] = spring[e] + phys[e];
        }
    }
    free(spring); free(phys); free(f_first); free(f_last);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* RANMAR, libs/random_mars.cpp:11-103 (F. James, CPC 60 (1990) 329)       */

int orc_ranmars_init(orc_ranmars* r, int seed) {
    int ij, kl, i, j, k, l, ii, jj, m;
    double s, t;
    if (seed <= 0 || seed > 900000000) return ORC_E_ARG;
    r->save = 0;
    r->second = 0.0;
    memset(r->u, 0, sizeof(r->u));
    ij = (seed - 1) / 30082;
    kl = (seed - 1) - 30082 * ij;
    i = (ij / 177) % 177 + 2;
    j = ij % 177 + 2;
    k = (kl / 169) % 178 + 1;
    l = kl % 169;
    for (ii = 1; ii <= 97; ii++) {
        s = 0.0;
        t = 0.5;
        for (jj = 1; jj <= 24; jj++) {
            m = ((i * j) % 179) * k % 179;
            i = j; j = k; k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s = s + t;
            t = 0.5 * t;
        }
        r->u[ii] = s;
    }
    r->c = 362436.0 / 16777216.0;
    r->cd = 7654321.0 / 16777216.0;
    r->cm = 16777213.0 / 16777216.0;
    r->i97 = 97;
    r->j97 = 33;
    orc_ranmars_uniform(r);
    return ORC_OK;
}

double orc_ranmars_uniform(orc_ranmars* r) {
    double uni = r->u[r->i97] - r->u[r->j97];
    if (uni < 0.0) uni += 1.0;
    r->u[r->i97] = uni;
    r->i97--;
    if (r->i97 == 0) r->i97 = 97;
    r->j97--;
    if (r->j97 == 0) r->j97 = 97;
    r->c -= r->cd;
    if (r->c < 0.0) r->c += r->cm;
    uni -= r->c;
    if (uni < 0.0) uni += 1.0;
    return uni;
}

double orc_ranmars_gaussian(orc_ranmars* r) {
    double first, v1, v2, rsq, fac;
    if (!r->save) {
        do {
            v1 = 2.0 * orc_ranmars_uniform(r) - 1.0;
            v2 = 2.0 * orc_ranmars_uniform(r) - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while ((rsq >= 1.0) || (rsq == 0.0));
        fac = sqrt(-2.0 * log(rsq) / rsq);
        r->second = v1 * fac;
        first = v2 * fac;
        r->save = 1;
    } else {
        first = r->second;
        r->save = 0;
    }
    return first;
}

/* ------------------------------------------------------------------------ */
/* mt19937 (Matsumoto & Nishimura 1998) and the libstdc++ distributions the */
/* reference draws its initial state with (simulation.cpp:118-126, 200-217) */

void orc_mt_seed(orc_mt19937* g, unsigned int seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253U * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (unsigned int)i;
    g->mti = 624;
}

unsigned int orc_mt_next(orc_mt19937* g) {
    if (g->mti >= 624) {
        for (int i = 0; i < 624; i++) {
            unsigned int y = (g->mt[i] & 0x80000000U) | (g->mt[(i + 1) % 624] & 0x7fffffffU);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1U) ? 0x9908b0dfU : 0U);
        }
        g->mti = 0;
    }
    unsigned int y = g->mt[g->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680U;
    y ^= (y << 15) & 0xefc60000U;
    y ^= (y >> 18);
    return y;
}

/* std::generate_canonical<double, 53> over a 32-bit engine: two draws */
double orc_mt_canonical(orc_mt19937* g) {
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum += (double)orc_mt_next(g) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

double orc_mt_uniform_real(orc_mt19937* g, double a, double b) {
    return (orc_mt_canonical(g) * (b - a)) + a;
}

/* std::normal_distribution (Marsaglia polar), a fresh object per draw so the
 * cached second value is always discarded (simulation.cpp:213-217). */
double orc_mt_normal_fresh(orc_mt19937* g, double mean, double stddev) {
    double x, y, r2;
    do {
        x = 2.0 * orc_mt_canonical(g) - 1.0;
        y = 2.0 * orc_mt_canonical(g) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2 * log(r2) / r2);
    double ret = y * mult;
    return ret * stddev + mean;
}

/* ------------------------------------------------------------------------ */
/* Normal modes, src/normal_modes.cpp:47-81                                 */

void orc_nm_matrices(int P, double* C, double* CT) {
    const double pi = 3.14159265358979323846;
    for (int b = 0; b < P; ++b) {
        double pref;
        double fund_freq = 2 * pi / P * b;
        double* row = C + (size_t)b * P;
        if (b == 0) {
            pref = 1 / sqrt((double)P);
            for (int i = 0; i < P; ++i) row[i] = pref;
        } else if (b < 0.5 * P) {
            pref = sqrt(2.0 / P);
            for (int i = 0; i < P; ++i) row[i] = pref * cos(fund_freq * i);
        } else if (b == 0.5 * P) {
            pref = 1 / sqrt((double)P);
            for (int i = 0; i < P; ++i) row[i] = pref * (i % 2 == 0 ? 1.0 : -1.0);
        } else {
            pref = sqrt(2.0 / P);
            for (int i = 0; i < P; ++i) row[i] = -pref * sin(fund_freq * i);
        }
        double* crow = CT + (size_t)b * P;
        pref = sqrt(2.0 / P);
        for (int i = 0; i < P; ++i) crow[i] = 0.0;
        crow[0] = 1 / sqrt((double)P);
        for (int i = 1; i < 0.5 * P; ++i) crow[i] = pref * cos(fund_freq * i);
        if (P % 2 == 0) crow[P / 2] = 1 / sqrt((double)P) * (b % 2 == 0 ? 1.0 : -1.0);
        for (int i = (int)ceil(0.5 * (P + 1)); i < P; ++i) crow[i] = -pref * sin(fund_freq * i);
    }
}

/* ------------------------------------------------------------------------ */
/* Whole-ring simulation                                                    */

struct orc_sim {
    orc_params p;
    derived d;
    int N, P, D;
    size_t nb;
    double *x, *pm, *f;              /* [P][N][D] */
    double *ext_forces, *spring_forces, *tmpf;
    double *C, *CT, *nm_a, *nm_b;    /* NM matrices and scratch [P] */
    double *pnm;                     /* [P][N][D] */
    orc_ranmars* mars;
    orc_mt19937* mt;
    double c1, c2;                   /* Langevin coefficients */
    double *nm_freq;                 /* NM propagator, per bead */
    /* Nose-Hoover chains, one thermostat object per bead (rank): chains of
     * nchains components; per bead 1 chain (nose_hoover), N (np) or N*D (np_dim) */
    int nh_groups, nh_ndof;          /* chains per bead, dof per chain */
    double nh_Q1, nh_Qi, nh_required;
    double *eta, *eta_dot, *eta_dot_dot;   /* [P][nh_groups][nchains] */
};

orc_sim* orc_sim_create(const orc_params* p) {
    if (p->natoms < 1 || p->nbeads < 1 || p->ndim < 1 || p->ndim > 8) return NULL;
    if (p->thermostat >= ORC_THERMO_NOSE_HOOVER && p->nchains < 2) return NULL;   /* see nh_chain_step */
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->p = *p;
    s->d = derive(p);
    s->N = p->natoms; s->P = p->nbeads; s->D = p->ndim;
    s->nb = (size_t)s->N * s->D;
    size_t tot = s->nb * s->P;
    s->x = (double*)calloc(tot, sizeof(double));
    s->pm = (double*)calloc(tot, sizeof(double));
    s->f = (double*)calloc(tot, sizeof(double));
    s->ext_forces = (double*)calloc(tot, sizeof(double));
    s->spring_forces = (double*)calloc(tot, sizeof(double));
    s->tmpf = (double*)calloc(tot, sizeof(double));
    s->pnm = (double*)calloc(tot, sizeof(double));
    s->C = (double*)calloc((size_t)s->P * s->P, sizeof(double));
    s->CT = (double*)calloc((size_t)s->P * s->P, sizeof(double));
    s->nm_a = (double*)calloc((size_t)s->P, sizeof(double));
    s->nm_b = (double*)calloc((size_t)s->P, sizeof(double));
    s->nm_freq = (double*)calloc((size_t)s->P, sizeof(double));
    orc_nm_matrices(s->P, s->C, s->CT);
    s->mars = (orc_ranmars*)calloc((size_t)s->P, sizeof(orc_ranmars));
    s->mt = (orc_mt19937*)calloc((size_t)s->P, sizeof(orc_mt19937));
    for (int b = 0; b < s->P; ++b) {
        orc_ranmars_init(&s->mars[b], (int)(p->seed + b));
        orc_mt_seed(&s->mt[b], p->seed + b);
    }
    /* langevin.cpp:10-13 */
    s->c1 = exp(-0.5 * p->gamma * p->dt);
    s->c2 = sqrt((1 - s->c1 * s->c1) * p->mass / s->d.thermo_beta);
    /* normal_modes_propagator.cpp:13 */
    const double pi = 3.14159265358979323846;
    for (int b = 0; b < s->P; ++b) s->nm_freq[b] = 2 * s->d.omega_p * sin(b * pi / s->P);
    /* nose_hoover.cpp:10-32, 152-163, 198-208 (hbar = 1) */
    if (p->thermostat >= ORC_THERMO_NOSE_HOOVER) {
        const int D = s->D, N = s->N, nc = p->nchains;
        s->nh_Qi = 1.0 * 1.0 * s->d.beta / s->P;
        if (p->thermostat == ORC_THERMO_NOSE_HOOVER) {
            s->nh_groups = 1; s->nh_ndof = D * N;
            s->nh_Q1 = D * N * s->nh_Qi;
            s->nh_required = D * N / s->d.thermo_beta;
        } else if (p->thermostat == ORC_THERMO_NOSE_HOOVER_NP) {
            s->nh_groups = N; s->nh_ndof = D;
            s->nh_Q1 = D * s->nh_Qi;
            s->nh_required = D / s->d.thermo_beta;
        } else {
            s->nh_groups = N * D; s->nh_ndof = 1;
            s->nh_Q1 = s->nh_Qi;
            s->nh_required = 1 / s->d.thermo_beta;
        }
        const size_t n = (size_t)s->P * s->nh_groups * (nc > 0 ? nc : 1);
        s->eta = (double*)calloc(n, sizeof(double));
        s->eta_dot = (double*)calloc(n, sizeof(double));
        s->eta_dot_dot = (double*)calloc(n, sizeof(double));
    }
    return s;
}

void orc_sim_destroy(orc_sim* s) {
    if (!s) return;
    free(s->x); free(s->pm); free(s->f); free(s->ext_forces); free(s->spring_forces);
    free(s->tmpf); free(s->pnm); free(s->C); free(s->CT); free(s->nm_a); free(s->nm_b);
    free(s->nm_freq); free(s->mars); free(s->mt);
    free(s->eta); free(s->eta_dot); free(s->eta_dot_dot); free(s);
}

/* simulation.cpp:581-603; beads summed in order (see file header) */
void orc_sim_zero_momentum(orc_sim* s) {
    const int N = s->N, D = s->D, P = s->P;
    double cm[8] = {0};
    for (int b = 0; b < P; ++b) {
        double per[8] = {0};
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) per[a] += s->pm[s->nb * b + i * D + a];
        for (int a = 0; a < D; ++a) per[a] /= (N * P);
        for (int a = 0; a < D; ++a) cm[a] += per[a];
    }
    for (int b = 0; b < P; ++b)
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) s->pm[s->nb * b + i * D + a] -= cm[a];
}

void orc_sim_init_random(orc_sim* s) {
    const int N = s->N, D = s->D;
    for (int b = 0; b < s->P; ++b) {
        orc_mt19937* g = &s->mt[b];
        double* x = s->x + s->nb * b;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a)
                x[i * D + a] = orc_mt_uniform_real(g, -0.5 * s->p.size, 0.5 * s->p.size);
        double* pm = s->pm + s->nb * b;
        const double sd = 1 / sqrt(s->d.thermo_beta * s->p.mass);
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a)
                pm[i * D + a] = s->p.mass * orc_mt_normal_fresh(g, 0.0, sd);
    }
    orc_sim_zero_momentum(s);
    memset(s->f, 0, sizeof(double) * s->nb * s->P);
    memset(s->ext_forces, 0, sizeof(double) * s->nb * s->P);
    memset(s->spring_forces, 0, sizeof(double) * s->nb * s->P);
}

void orc_sim_set_state(orc_sim* s, const double* x, const double* p) {
    size_t tot = s->nb * s->P;
    if (x) memcpy(s->x, x, sizeof(double) * tot);
    if (p) memcpy(s->pm, p, sizeof(double) * tot);
}

void orc_sim_get_state(const orc_sim* s, double* x, double* p, double* f) {
    size_t tot = s->nb * s->P;
    if (x) memcpy(x, s->x, sizeof(double) * tot);
    if (p) memcpy(p, s->pm, sizeof(double) * tot);
    if (f) memcpy(f, s->f, sizeof(double) * tot);
}

void orc_sim_draw_noise(orc_sim* s, double* noise) {
    for (int o = 0; o < 2; ++o)
        for (int b = 0; b < s->P; ++b)
            for (size_t e = 0; e < s->nb; ++e)
                noise[(size_t)o * s->nb * s->P + s->nb * b + e] = orc_ranmars_gaussian(&s->mars[b]);
}

int orc_sim_forces(orc_sim* s) {
    return orc_update_forces(&s->p, s->x, s->f);
}

/* thermostat.cpp:15-19 + langevin.cpp:15-27 + thermostat_coupling.cpp */
/* NoseHooverThermostat::singleChainStep, nose_hoover.cpp:96-150 (Tuckerman
 * et al. 2006, LAMMPS), on one chain (e, ed, edd: its nchains components). */
static double nh_chain_step(const orc_sim* s, double* e, double* ed, double* edd, double current_energy) {
    const int nc = s->p.nchains;
    const double dt2 = 0.5 * s->p.dt, dt4 = 0.25 * s->p.dt, dt8 = 0.125 * s->p.dt;
    const double Q1 = s->nh_Q1, Qi = s->nh_Qi, tb = s->d.thermo_beta;
    double exp_factor = 0.0;
    edd[0] = (current_energy - s->nh_required) / Q1;
    ed[nc - 1] += edd[nc - 1] * dt4;
    for (int i = nc - 2; i >= 0; i--) {
        exp_factor = exp(-dt8 * ed[i + 1]);
        ed[i] *= exp_factor;
        ed[i] += edd[i] * dt4;
        ed[i] *= exp_factor;
    }
    double scale = exp(-dt2 * ed[0]);
    for (int i = 0; i < nc; i++) e[i] += dt2 * ed[i];
    edd[0] = (current_energy * scale * scale - s->nh_required) / Q1;
    ed[0] *= exp_factor;
    ed[0] += edd[0] * dt4;
    ed[0] *= exp_factor;
    double Q_former = Q1;
    for (int i = 1; i < nc - 1; i++) {
        exp_factor = exp(-dt8 * ed[i + 1]);
        ed[i] *= exp_factor;
        edd[i] = (Q_former * ed[i - 1] * ed[i - 1] - 1 / tb) / Qi;
        ed[i] += edd[i] * dt4;
        ed[i] *= exp_factor;
        Q_former = Qi;
    }
    edd[nc - 1] = (Qi * ed[nc - 2] * ed[nc - 2] - 1 / tb) / Qi;
    ed[nc - 1] += edd[nc - 1] * dt4;
    return scale;
}

/* momentaUpdate of the three Nose-Hoover variants (nose_hoover.cpp:75-94,
 * 165-184, 210-227) on bead b's momenta pb (Cartesian, or its normal mode) */
static void nh_bead(orc_sim* s, int b, double* pb) {
    const int N = s->N, D = s->D, nc = s->p.nchains;
    const size_t base = (size_t)b * s->nh_groups * nc;
    if (s->p.thermostat == ORC_THERMO_NOSE_HOOVER) {
        double ce = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) ce += pb[i * D + a] * pb[i * D + a];
        ce /= s->p.mass;
        const double sc = nh_chain_step(s, s->eta + base, s->eta_dot + base, s->eta_dot_dot + base, ce);
        for (int e = 0; e < N * D; ++e) pb[e] = pb[e] * sc;
    } else if (s->p.thermostat == ORC_THERMO_NOSE_HOOVER_NP) {
        for (int i = 0; i < N; ++i) {
            double ce = 0.0;
            for (int a = 0; a < D; ++a) ce += pb[i * D + a] * pb[i * D + a];
            ce /= s->p.mass;
            const size_t o = base + (size_t)i * nc;
            const double sc = nh_chain_step(s, s->eta + o, s->eta_dot + o, s->eta_dot_dot + o, ce);
            for (int a = 0; a < D; ++a) pb[i * D + a] = pb[i * D + a] * sc;
        }
    } else {
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) {
                const double ce = pb[i * D + a] * pb[i * D + a] / s->p.mass;
                const size_t o = base + ((size_t)i * D + a) * nc;
                const double sc = nh_chain_step(s, s->eta + o, s->eta_dot + o, s->eta_dot_dot + o, ce);
                pb[i * D + a] = pb[i * D + a] * sc;
            }
    }
}

/* getAdditionToH summed over the beads' thermostats (nose_hoover.cpp:41-66,
 * 186-192, 229-237), the nh_energy column (classical.cpp:24-26) */
static double nh_addition_to_h(const orc_sim* s) {
    const int nc = s->p.nchains;
    const double tb = s->d.thermo_beta;
    double tot = 0.0;
    for (int b = 0; b < s->P; ++b) {
        double add = 0.0;
        for (int g = 0; g < s->nh_groups; ++g) {
            const size_t o = ((size_t)b * s->nh_groups + g) * nc;
            double a = 0.5 * s->nh_Q1 * s->eta_dot[o] * s->eta_dot[o];
            a += s->nh_ndof * s->eta[o] / tb;
            for (int i = 1; i < nc; i++) {
                a += 0.5 * s->nh_Qi * s->eta_dot[o + i] * s->eta_dot[o + i];
                a += s->eta[o + i] / tb;
            }
            add += a;
        }
        tot += add;
    }
    return tot;
}

static void thermostat_step(orc_sim* s, const double* noise) {
    if (s->p.thermostat >= ORC_THERMO_NOSE_HOOVER) {
        const int N = s->N, D = s->D, P = s->P;
        if (!s->p.nmthermostat) {
            for (int b = 0; b < P; ++b) nh_bead(s, b, s->pm + s->nb * b);
            return;
        }
        /* NM coupling (thermostat_coupling.cpp:28-47): mode b from the shared
         * cartesian momenta, thermostatted, then all beads rebuilt */
        for (int b = 0; b < P; ++b) {
            const double* row = s->C + (size_t)b * P;
            for (int i = 0; i < N; ++i)
                for (int a = 0; a < D; ++a) {
                    const size_t off = (size_t)i * D + a;
                    double pnm = 0;
                    for (int j = 0; j < P; ++j) pnm += row[j] * s->pm[s->nb * j + off];
                    s->pnm[s->nb * b + off] = pnm;
                }
            nh_bead(s, b, s->pnm + s->nb * b);
        }
        for (int b = 0; b < P; ++b) {
            const double* crow = s->CT + (size_t)b * P;
            for (size_t off = 0; off < s->nb; ++off) {
                double pc = 0;
                for (int k = 0; k < P; ++k) pc += crow[k] * s->pnm[s->nb * k + off];
                s->tmpf[s->nb * b + off] = pc;
            }
        }
        memcpy(s->pm, s->tmpf, sizeof(double) * s->nb * P);
        return;
    }
    if (s->p.thermostat != ORC_THERMO_LANGEVIN) return;   /* "none": no-op */
    const int N = s->N, D = s->D, P = s->P;
    if (!s->p.nmthermostat) {
        for (int b = 0; b < P; ++b)
            for (int i = 0; i < N; ++i)
                for (int a = 0; a < D; ++a) {
                    size_t e = s->nb * b + i * D + a;
                    double xi = noise ? noise[e] : orc_ranmars_gaussian(&s->mars[b]);
                    s->pm[e] = s->c1 * s->pm[e] + s->c2 * xi;
                }
        return;
    }
    /* NM coupling: each rank b updates mode b from the shared cartesian
     * momenta, then every rank rebuilds its cartesian momentum. */
    for (int b = 0; b < P; ++b) {
        const double* row = s->C + (size_t)b * P;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) {
                size_t off = (size_t)i * D + a;
                double xi = noise ? noise[s->nb * b + off] : orc_ranmars_gaussian(&s->mars[b]);
                double pnm = 0;
                for (int j = 0; j < P; ++j) pnm += row[j] * s->pm[s->nb * j + off];
                s->pnm[s->nb * b + off] = s->c1 * pnm + s->c2 * xi;
            }
    }
    for (int b = 0; b < P; ++b) {
        const double* crow = s->CT + (size_t)b * P;
        for (size_t off = 0; off < s->nb; ++off) {
            double pc = 0;
            for (int k = 0; k < P; ++k) pc += crow[k] * s->pnm[s->nb * k + off];
            s->tmpf[s->nb * b + off] = pc;
        }
    }
    memcpy(s->pm, s->tmpf, sizeof(double) * s->nb * P);
}

/* propagator: velocity_verlet.cpp:7-38 */
static int vv_step(orc_sim* s) {
    const size_t tot = s->nb * s->P;
    for (size_t e = 0; e < tot; ++e) s->pm[e] += 0.5 * s->p.dt * s->f[e];
    for (size_t e = 0; e < tot; ++e) s->x[e] += s->p.dt * s->pm[e] / s->p.mass;
    int rc = orc_sim_forces(s);
    for (size_t e = 0; e < tot; ++e) s->pm[e] += 0.5 * s->p.dt * s->f[e];
    return rc;
}

/* normal_modes_propagator.cpp:73-103, distinguishable branch only */
static void nm_external_kick(orc_sim* s) {
    const size_t tot = s->nb * s->P;
    for (size_t e = 0; e < tot; ++e) s->pm[e] += 0.5 * s->p.dt * s->ext_forces[e];
}

/* normal_modes_propagator.cpp:19-71 */
static int nm_step(orc_sim* s) {
    const int P = s->P, D = s->D, N = s->N;
    const size_t nb = s->nb;
    nm_external_kick(s);
    double* qn = s->tmpf;   /* NM coords [P][N][D] */
    for (int b = 0; b < P; ++b) {
        const double* row = s->C + (size_t)b * P;
        const double freq = s->nm_freq[b];
        const double c = cos(freq * s->p.dt), sn = sin(freq * s->p.dt);
        const double m_omega = s->p.mass * freq;
        for (size_t off = 0; off < nb; ++off) {
            double q = 0, pp = 0;
            for (int j = 0; j < P; ++j) q += row[j] * s->x[nb * j + off];
            for (int j = 0; j < P; ++j) pp += row[j] * s->pm[nb * j + off];
            if (freq == 0) {
                qn[nb * b + off] = q + s->p.dt / s->p.mass * pp;
                s->pnm[nb * b + off] = pp;
            } else {
                qn[nb * b + off] = c * q + sn / m_omega * pp;
                s->pnm[nb * b + off] = (-1) * m_omega * sn * q + c * pp;
            }
        }
    }
    (void)N; (void)D;
    /* back to cartesian (each rank its own bead) */
    double* xc = (double*)malloc(sizeof(double) * nb * P);
    double* pc = (double*)malloc(sizeof(double) * nb * P);
    for (int b = 0; b < P; ++b) {
        const double* crow = s->CT + (size_t)b * P;
        for (size_t off = 0; off < nb; ++off) {
            double xx = 0, pv = 0;
            for (int k = 0; k < P; ++k) xx += crow[k] * qn[nb * k + off];
            for (int k = 0; k < P; ++k) pv += crow[k] * s->pnm[nb * k + off];
            xc[nb * b + off] = xx;
            pc[nb * b + off] = pv;
        }
    }
    memcpy(s->x, xc, sizeof(double) * nb * P);
    memcpy(s->pm, pc, sizeof(double) * nb * P);
    free(xc); free(pc);
    int rc = orc_sim_forces(s);
    /* updatePhysicalForces(ext_forces) per bead */
    for (int b = 0; b < P; ++b) orc_physical_forces(&s->p, s->x + nb * b, s->ext_forces + nb * b);
    nm_external_kick(s);
    return rc;
}

int orc_sim_step(orc_sim* s, const double* noise) {
    const size_t tot = s->nb * s->P;
    thermostat_step(s, noise);
    if (s->p.fixcom) orc_sim_zero_momentum(s);
    int rc = (s->p.propagator == ORC_PROP_NORMAL_MODES && !is_bosonic(&s->p)) ? nm_step(s) : vv_step(s);
    thermostat_step(s, noise ? noise + tot : NULL);
    if (s->p.fixcom) orc_sim_zero_momentum(s);
    return rc;
}

/* observables: energy.cpp:21-111, classical.cpp:131-188, bosonic.cpp:17-22 */
int orc_sim_observables(orc_sim* s, double* out) {
    const int N = s->N, D = s->D, P = s->P;
    const size_t nb = s->nb;
    const orc_params* p = &s->p;
    double kinetic = 0, potential = 0, ext_pot_t = 0, int_pot_t = 0, virial_t = 0;
    double temperature = 0, cl_kin = 0, cl_spring = 0, prob_dist = 0, prob_all = 0;
    const double cutoff = orc_effective_cutoff(p);
    double *E = NULL, *V = NULL;
    int rc = ORC_OK;
    const int factorial = is_bosonic(p) && p->exchange_algo == ORC_EXCHANGE_FACTORIAL;
    if (factorial && N > ORC_FACTORIAL_MAXN) return ORC_E_ARG;
    if (is_bosonic(p) && !factorial) {
        E = (double*)malloc(sizeof(double) * (size_t)N * (N + 1) / 2);
        V = (double*)malloc(sizeof(double) * (N + 1));
        double* Vb = (double*)malloc(sizeof(double) * (N + 1));
        orc_cycle_energies(N, D, s->x, s->x + nb * (P - 1), s->d.k, p->pbc, p->size, E);
        rc = orc_v_forward(N, s->d.beta_t, E, V);
        if (rc == ORC_OK) rc = orc_v_backward(N, s->d.beta_t, E, V, Vb);
        free(Vb);
        if (rc != ORC_OK) { free(E); free(V); return rc; }
    }
    for (int b = 0; b < P; ++b) {
        const double* coord = s->x + nb * b;
        const double* prev = s->x + nb * ((b - 1 + P) % P);
        const double* mom = s->pm + nb * b;
        /* classicalSpringEnergy, simulation.cpp:464-486 */
        double spring_b = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) {
                double diff = prev[i * D + a] - coord[i * D + a];
                if (p->pbc) diff = orc_min_image(diff, p->size);
                spring_b += diff * diff;
            }
        spring_b *= 0.5 * s->d.k;
        double kin = 0.5 * D * N / s->d.beta;
        if (b == 0 && factorial) {
            const double* xP = s->x + nb * (P - 1);
            const double e_shift = orc_factorial_e_shift(N, D, s->d.k, p->pbc, p->size, s->x, xP);
            kin += orc_factorial_prim_estimator(N, P, D, s->d.beta_t, s->d.k, p->pbc, p->size, s->x, xP, e_shift);
        } else if (b == 0 && is_bosonic(p)) kin += orc_prim_estimator(N, P, s->d.beta_t, E, V);
        else kin -= spring_b / P;
        kinetic += kin;
        double pot = 0, virial = 0, int_pot = 0, ext_pot = 0;
        if (p->ext_kind != ORC_POT_FREE) {
            double g[8];
            ext_pot = orc_ext_V(p, coord);
            pot += ext_pot;
            for (int i = 0; i < N; ++i) {
                orc_ext_grad(p, coord + i * D, g);
                for (int a = 0; a < D; ++a) virial -= coord[i * D + a] * (g[a] * (-1.0));
            }
        }
        if (cutoff != 0.0) {
            for (int i = 0; i < N; ++i)
                for (int j = i + 1; j < N; ++j) {
                    double diff[8], g[8];
                    pair_sep(p, coord, i, j, diff);
                    const double distance = vnorm(D, diff);
                    if (distance < cutoff || cutoff < 0.0) {
                        int_grad(p, diff, g);
                        double v = int_V(p, diff);
                        pot += v;
                        int_pot += v;
                        for (int a = 0; a < D; ++a) virial -= coord[i * D + a] * (g[a] * (-1.0));
                    }
                }
        }
        if (p->ext_kind != ORC_POT_FREE && p->int_kind != ORC_POT_FREE) {
            ext_pot /= P;
            int_pot /= P;
            ext_pot_t += ext_pot;
            int_pot_t += int_pot;
        }
        if (p->ext_kind != ORC_POT_FREE || p->int_kind != ORC_POT_FREE) {
            pot /= P;
            virial *= 0.5 / P;
            potential += pot;
            virial_t += virial;
        }
        /* classical.cpp:142-171 */
        double ke = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) ke += mom[i * D + a] * mom[i * D + a];
        ke *= 0.5 / p->mass;
        cl_kin += ke;
        double dof = (double)D * N * P;
        double temp_b = 2.0 * ke / (dof * 1.0);
        temp_b /= P;
        temperature += temp_b;
        if (b == 0 && factorial)
            cl_spring += orc_factorial_vb(N, D, s->d.beta_t, s->d.k, p->pbc, p->size, s->x, s->x + nb * (P - 1));
        else if (b == 0 && is_bosonic(p)) cl_spring += V[N];
        else cl_spring += spring_b;
    }
    /* the factorial exchange returns 0 for both (factorial_bosonic_exchange.cpp:237-253) */
    if (is_bosonic(p) && !factorial) {
        prob_dist = orc_distinct_probability(N, s->d.beta_t, E, V);
        prob_all = orc_longest_probability(N, s->d.beta_t, E, V);
    }
    out[0] = kinetic; out[1] = potential; out[2] = ext_pot_t; out[3] = int_pot_t;
    out[4] = virial_t; out[5] = temperature; out[6] = cl_kin; out[7] = cl_spring;
    out[8] = prob_dist; out[9] = prob_all;
    out[10] = p->thermostat >= ORC_THERMO_NOSE_HOOVER ? nh_addition_to_h(s) : 0.0;
    free(E); free(V);
    return rc;
}
