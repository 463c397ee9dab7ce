/*
 * pimdb_oracle_quad.c -- the Feldman-Hirshberg exchange of pimdb_oracle.c in
 * quadruple precision (__float128, libquadmath): the exact value, to about
 * 1e-30 relative, of the reference's exchange for given double inputs.
 *
 * TEST INFRASTRUCTURE ONLY (see pimdb_oracle.h).  It measures how far the
 * double-precision results (the oracle's, which follow the reference's
 * arithmetic order, and the GPU's) lie from the exact ones where the problem
 * is ill-conditioned (N = 1024).  Same steps as orc_exchange, each citing the
 * reference:
 *   cycle energies     quadratic_bosonic_exchange.cpp:34-61 (the recurrence)
 *   forward V          :73-99
 *   backward Vb        :101-128
 *   probabilities      :142-157
 *   exterior forces    :159-215 (with the interior links, as orc_exchange)
 * Pinned by: at N <= 8 it agrees with the double oracle to its rounding and
 * with orc_vb_enumerate; its forces equal -dV_B/dx by central differences of
 * its own V_B (tests/test_oracle_math.py).
 */
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

#include "pimdb_oracle.h"

typedef __float128 qd;

/* src/common.cpp:41-43 in quad */
static qd q_min_image(qd dx, qd L) { return dx - L * floorq(dx / L + (qd)0.5); }

/* bosonic_exchange_base.cpp:30-64: |x2[l2] - x1[l1]|^2 */
static qd q_sep_sq(int N, int D, const double* x1, int l1, const double* x2, int l2, int pbc, qd L) {
    l1 %= N;
    l2 %= N;
    qd s = 0;
    for (int a = 0; a < D; ++a) {
        qd d = (qd)x2[l2 * D + a] - (qd)x1[l1 * D + a];
        if (pbc) d = q_min_image(d, L);
        s += d * d;
    }
    return s;
}

static void q_sep(int N, int D, const double* x1, int l1, const double* x2, int l2, int pbc, qd L, qd* d) {
    l1 %= N;
    l2 %= N;
    for (int a = 0; a < D; ++a) {
        qd v = (qd)x2[l2 * D + a] - (qd)x1[l1 * D + a];
        d[a] = pbc ? q_min_image(v, L) : v;
    }
}

static inline size_t enk(int m, int k) { return (size_t)m * (m + 1) / 2 - k; }

int orc_exchange_quad(int N, int D, double beta_t, double kd, int pbc, double Ld, const double* x1, const double* x2,
                      const double* xPm1, const double* xP, double* V_out, double* Vb_out, double* prob_out,
                      double* f_first, double* f_last) {
    if (N < 1 || D < 1 || D > 8) return ORC_E_ARG;
    const qd beta = beta_t, k = kd, L = Ld, half = (qd)0.5;
    const size_t ne = (size_t)N * (N + 1) / 2;
    qd* E = (qd*)malloc(sizeof(qd) * ne);
    qd* V = (qd*)malloc(sizeof(qd) * (N + 1));
    qd* Vb = (qd*)malloc(sizeof(qd) * (N + 1));
    qd* tmp = (qd*)malloc(sizeof(qd) * (N + 1));
    qd* pr = (qd*)malloc(sizeof(qd) * (size_t)N * N);
    if (!E || !V || !Vb || !tmp || !pr) {
        free(E); free(V); free(Vb); free(tmp); free(pr);
        return ORC_E_ARG;
    }
    /* :34-61 with x_first = bead 1, x_last = bead P */
    for (int v = 0; v < N; ++v) {
        E[enk(v + 1, 1)] = half * k * q_sep_sq(N, D, x1, v, xP, v, pbc, L);
        for (int u = v - 1; u >= 0; --u)
            E[enk(v + 1, v - u + 1)] = E[enk(v + 1, v - u)] +
                                       half * k * (q_sep_sq(N, D, xP, u, x1, u + 1, pbc, L) -
                                                   q_sep_sq(N, D, x1, u + 1, xP, v, pbc, L) +
                                                   q_sep_sq(N, D, x1, u, xP, v, pbc, L));
    }
    /* :73-99 */
    V[0] = 0;
    for (int m = 1; m <= N; ++m) {
        qd sh = FLT128_MAX;
        for (int kk = m; kk > 0; --kk) {
            tmp[kk - 1] = E[enk(m, kk)] + V[m - kk];
            if (tmp[kk - 1] < sh) sh = tmp[kk - 1];
        }
        qd s = 0;
        for (int kk = m; kk > 0; --kk) s += expq(-beta * (tmp[kk - 1] - sh));
        V[m] = sh - logq(s / (qd)m) / beta;
    }
    /* :101-128 */
    Vb[N] = 0;
    for (int l = N - 1; l > 0; --l) {
        qd sh = FLT128_MAX;
        for (int p = l; p < N; ++p) {
            tmp[p] = E[enk(p + 1, p - l + 1)] + Vb[p + 1];
            if (tmp[p] < sh) sh = tmp[p];
        }
        qd s = 0;
        for (int p = l; p < N; ++p) s += expq(-beta * (tmp[p] - sh)) / (qd)(p + 1);
        Vb[l] = sh - logq(s) / beta;
    }
    Vb[0] = V[N];
    /* :142-157 */
    memset(pr, 0, sizeof(qd) * (size_t)N * N);
    for (int l = 0; l < N - 1; ++l) pr[(size_t)N * l + l + 1] = 1 - expq(-beta * (V[l + 1] + Vb[l + 1] - V[N]));
    for (int u = 0; u < N; ++u)
        for (int l = u; l < N; ++l)
            pr[(size_t)N * l + u] = expq(-beta * (V[u] + E[enk(l + 1, l - u + 1)] + Vb[l + 1] - V[N])) / (qd)(l + 1);
    /* :188-215 (bead 1) and :159-186 (bead P) */
    qd d[8], s[8];
    for (int l = 0; l < N && f_first; ++l) {
        for (int a = 0; a < D; ++a) s[a] = 0;
        for (int j = (l - 1 > 0 ? l - 1 : 0); j < N; ++j) {
            q_sep(N, D, x1, l, xP, j, pbc, L, d);
            for (int a = 0; a < D; ++a) s[a] += pr[(size_t)N * j + l] * d[a];
        }
        q_sep(N, D, x1, l, x2, l, pbc, L, d);
        for (int a = 0; a < D; ++a) f_first[l * D + a] = (double)((s[a] + d[a]) * k);
    }
    for (int l = 0; l < N && f_last; ++l) {
        for (int a = 0; a < D; ++a) s[a] = 0;
        for (int j = 0; j <= l + 1 && j < N; ++j) {
            q_sep(N, D, xP, l, x1, j, pbc, L, d);
            for (int a = 0; a < D; ++a) s[a] += pr[(size_t)N * l + j] * d[a];
        }
        q_sep(N, D, xP, l, xPm1, l, pbc, L, d);
        for (int a = 0; a < D; ++a) f_last[l * D + a] = (double)((s[a] + d[a]) * k);
    }
    for (int i = 0; i <= N; ++i) {
        if (V_out) V_out[i] = (double)V[i];
        if (Vb_out) Vb_out[i] = (double)Vb[i];
    }
    if (prob_out)
        for (size_t i = 0; i < (size_t)N * N; ++i) prob_out[i] = (double)pr[i];
    free(E); free(V); free(Vb); free(tmp); free(pr);
    return ORC_OK;
}
