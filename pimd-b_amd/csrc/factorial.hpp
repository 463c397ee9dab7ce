// factorial.hpp -- the factorial bosonic exchange on the GPU: the effective
// spring potential of beads 1 and P averaged over all N! permutations, its
// forces on beads 1 and P and its kinetic-estimator term.
//
// Reference: src/bosonic_exchange/factorial_bosonic_exchange.cpp (the
// reference's compile-time alternative to the Feldman-Hirshberg recursion,
// FACTORIAL_BOSONIC_ALGORITHM).  For a permutation sigma the last bead of
// particle m connects to the first bead of particle sigma(m), with
//     E_sigma = k/2 sum_m |d(sigma(m), m)|^2,   d(a, m) = x1[a] - xP[m] (min image),
// and weight w_sigma = exp(-beta~ (E_sigma - e_shift)), e_shift = min E_sigma
// (getMinExteriorSpringEnergy, :57-79).  The forces (:121-230) are averages
// over sigma, so they only need the marginal weights
//     W[a][m] = sum_{sigma: sigma(m) = a} w_sigma,   S = sum_sigma w_sigma:
//     F_P[m] = k (xPm1[m] - xP[m] + sum_a W[a][m] d(a, m) / S)
//     F_1[a] = k (x2[a] - x1[a]   - sum_m W[a][m] d(a, m) / S)
// and the estimator (:258-290) is -(sum_sigma E_sigma w_sigma / S) / P.  The
// effective potential (:88-114) is V = e_shift - ln(S / N!) / beta~, the
// reference's -ln(mean exp(-beta~ E)) / beta~ with the shift taken out.
//
// One wave enumerates the permutations: lane t takes a contiguous range of
// lexicographic ranks, builds its first permutation from the rank (Lehmer
// code) and walks the rest with next_permutation, in LDS.  It keeps its own
// W in LDS; the 64 partial W are added in lane order, so the result depends
// on the data only.  E_sigma is summed over m in increasing order, as the
// reference sums over its particle loop.
#pragma once

#include "device_common.hpp"

namespace pimdb {

constexpr int FACT_MAXN = 10;    // 10! = 3628800 permutations
constexpr int FACT_LANES = 64;   // one wave

__host__ __device__ inline double factorial_d(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

// LDS doubles of k_factorial_exchange (the int permutation arrays are
// counted as doubles, rounded up).
__host__ __device__ inline size_t factorial_lds_doubles(int N, int D) {
    return (size_t)N * N * D + (size_t)N * N + (size_t)FACT_LANES * N * N + ((size_t)FACT_LANES * N + 1) / 2 + 8;
}

// std::next_permutation on p[0..n) (lexicographic order); false after the last.
__device__ __forceinline__ bool fact_next_perm(int* p, int n) {
    int i = n - 2;
    while (i >= 0 && p[i] >= p[i + 1]) --i;
    if (i < 0) return false;
    int j = n - 1;
    while (p[j] <= p[i]) --j;
    int t = p[i];
    p[i] = p[j];
    p[j] = t;
    for (int lo = i + 1, hi = n - 1; lo < hi; ++lo, --hi) {
        t = p[lo];
        p[lo] = p[hi];
        p[hi] = t;
    }
    return true;
}

// The permutation of lexicographic rank r.
__device__ __forceinline__ void fact_unrank(long long r, int n, int* p) {
    unsigned used = 0;
    for (int i = 0; i < n; ++i) {
        long long f = 1;
        for (int q = 2; q <= n - 1 - i; ++q) f *= q;
        long long c = r / f;
        r %= f;
        for (int v = 0; v < n; ++v) {
            if (used >> v & 1u) continue;
            if (c == 0) {
                p[i] = v;
                used |= 1u << v;
                break;
            }
            --c;
        }
    }
}

// x1, x2, xPm1, xP: beads 1, 2, P-1, P as SoA [a][i].  Outputs (each may be
// null): f1 / fP += the exchange forces on beads 1 and P (added to what they
// hold, as the plain form adds the physical forces), fex[2][D][N] the forces
// alone, VN = V_eff, est[0..2] = kinetic term, distinct and longest
// probabilities (the reference returns 0 for both, :237-253).  Launch: one
// block of FACT_LANES threads, factorial_lds_doubles(N, D) doubles of LDS.
template <int D>
static __global__ void __launch_bounds__(FACT_LANES) k_factorial_exchange(
    int N, int P, double bt, double k, int pbc, double L, const double* __restrict__ x1,
    const double* __restrict__ x2, const double* __restrict__ xPm1, const double* __restrict__ xP,
    double* __restrict__ f1, double* __restrict__ fP, double* __restrict__ fex, double* __restrict__ VN,
    double* __restrict__ est) {
    extern __shared__ double smem[];
    double* dv = smem;                              // [N][N][D]   d(a, m)
    double* D2 = dv + (size_t)N * N * D;            // [N][N]      |d(a, m)|^2
    double* Wp = D2 + (size_t)N * N;                // [lane][N][N] partial W
    int* perm = (int*)(Wp + (size_t)FACT_LANES * N * N);   // [lane][N]
    const int t = threadIdx.x;
    for (int e = t; e < N * N; e += FACT_LANES) {
        const int a = e / N, m = e % N;
        double d[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            d[c] = disp(xP[c * N + m], x1[c * N + a], pbc != 0, L);
            dv[(size_t)e * D + c] = d[c];
        }
        D2[e] = norm2<D>(d);
    }
    for (int e = t; e < FACT_LANES * N * N; e += FACT_LANES) Wp[e] = 0.0;
    __syncthreads();
    const long long nperm = (long long)(factorial_d(N) + 0.5);
    const long long chunk = (nperm + FACT_LANES - 1) / FACT_LANES;
    const long long r0 = min(nperm, chunk * t), r1 = min(nperm, r0 + chunk);
    int* p = perm + (size_t)t * N;
    // pass 1: the smallest exterior spring energy (e_shift)
    double emin = INFINITY;
    if (r0 < r1) {
        fact_unrank(r0, N, p);
        for (long long r = r0; r < r1; ++r) {
            double s = 0.0;
            for (int m = 0; m < N; ++m) s += D2[p[m] * N + m];
            emin = fmin(emin, s);
            if (r + 1 < r1) fact_next_perm(p, N);
        }
    }
    emin = wave_min(emin);
    const double e_shift = 0.5 * k * emin;
    // pass 2: weights, marginal weights, estimator sum
    double S = 0.0, SE = 0.0;
    double* W = Wp + (size_t)t * N * N;
    if (r0 < r1) {
        fact_unrank(r0, N, p);
        for (long long r = r0; r < r1; ++r) {
            double s = 0.0;
            for (int m = 0; m < N; ++m) s += D2[p[m] * N + m];
            const double e = 0.5 * k * s;
            const double w = exp(-bt * (e - e_shift));
            S += w;
            SE += e * w;
            for (int m = 0; m < N; ++m) W[p[m] * N + m] += w;
            if (r + 1 < r1) fact_next_perm(p, N);
        }
    }
    S = wave_sum(S);
    SE = wave_sum(SE);
    __syncthreads();
    // W[a][m] summed over lanes in lane order; each thread owns entries
    for (int e = t; e < N * N; e += FACT_LANES) {
        double w = 0.0;
        for (int q = 0; q < FACT_LANES; ++q) w += Wp[(size_t)q * N * N + e];
        Wp[e] = w;   // lane 0's slot holds the total (its own part is already read)
    }
    __syncthreads();
    const double invS = 1.0 / S;
    for (int l = t; l < N; l += FACT_LANES) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double aP = 0.0, a1 = 0.0;
            for (int q = 0; q < N; ++q) {
                aP += Wp[q * N + l] * dv[((size_t)q * N + l) * D + c];   // sum_a W[a][l] d(a, l)
                a1 += Wp[l * N + q] * dv[((size_t)l * N + q) * D + c];   // sum_m W[l][m] d(l, m)
            }
            const double FP = k * (disp(xP[c * N + l], xPm1[c * N + l], pbc != 0, L) + aP * invS);
            const double F1 = k * (disp(x1[c * N + l], x2[c * N + l], pbc != 0, L) - a1 * invS);
            if (fex) {
                fex[(size_t)c * N + l] = F1;
                fex[((size_t)D + c) * N + l] = FP;
            }
            if (f1) f1[c * N + l] = F1 + f1[c * N + l];
            if (fP) fP[c * N + l] = FP + fP[c * N + l];
        }
    }
    if (t == 0) {
        if (VN) *VN = e_shift - log(S / (double)nperm) / bt;
        if (est) {
            est[0] = -(SE * invS) / P;
            est[1] = 0.0;
            est[2] = 0.0;
        }
    }
}

}  // namespace pimdb
