// observables.hpp -- per-bead sums and the bosonic estimators behind the
// energy, classical and bosonic observables.
//
// Reference: EnergyObservable (src/observables/energy.cpp:21-111),
// ClassicalObservable (src/observables/classical.cpp:131-188),
// BosonicObservable (src/observables/bosonic.cpp:17-22),
// Simulation::classicalSpringEnergy (src/simulation.cpp:464-486),
// BosonicExchange::primEstimator / getDistinctProbability /
// getLongestProbability (quadratic_bosonic_exchange.cpp:222-279).
#pragma once

#include "device_common.hpp"

namespace pimdb {

// out[b] = {classical spring energy of link (b-1, b), sum of p^2}.
// One workgroup per bead.
template <int D>
static __global__ void __launch_bounds__(256) k_bead_sums(int N, int P, double k, int pbc, double L,
                                                   const double* __restrict__ x, const double* __restrict__ p,
                                                   double* __restrict__ out) {
    __shared__ double red[2][256 / WAVE];
    const int b = blockIdx.x;
    const size_t bs = (size_t)D * N;
    const double* xb = x + bs * b;
    const double* xprev = x + bs * ((b - 1 + P) % P);
    const double* pb = p + bs * b;
    double sp = 0.0, kp = 0.0;
    for (int e = threadIdx.x; e < D * N; e += blockDim.x) {
        const double d = disp(xb[e], xprev[e], pbc != 0, L);
        sp += d * d;
        kp += pb[e] * pb[e];
    }
    sp = wave_sum(sp);
    kp = wave_sum(kp);
    if (threadIdx.x % WAVE == 0) {
        red[0][threadIdx.x / WAVE] = sp;
        red[1][threadIdx.x / WAVE] = kp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < (int)blockDim.x / WAVE; ++w) { s0 += red[0][w]; s1 += red[1][w]; }
        out[2 * b] = s0 * (0.5 * k);
        out[2 * b + 1] = s1;
    }
}

// Primitive kinetic estimator of the exterior springs and the topology
// probabilities from E and V (forward pass done).  One workgroup.
//   prim[m] = sum_{j<m} w(m,j) (prim[j] - ecyc(j, m-1)),
//   w(m,j)  = exp(-bt (ecyc(j,m-1) + V[j] - V[m])) / m
// which is the reference's shifted form (:250-279) with its common shift
// cancelled.  out = {prim[N]/P, prob_dist, prob_all}.
static __global__ void __launch_bounds__(1024) k_prim_estimator(int N, int P, double bt, const double* __restrict__ E,
                                                         const double* __restrict__ V, double* __restrict__ out) {
    extern __shared__ double smem[];
    double* prim = smem;                 // [N+1]
    double* red = smem + N + 1;          // [nwave]
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE, nwave = blockDim.x / WAVE;
    if (tid == 0) prim[0] = 0.0;
    __syncthreads();
    for (int m = 1; m <= N; ++m) {
        const size_t col = (size_t)(m - 1) * m / 2;
        double s = 0.0;
        for (int j = tid; j < m; j += blockDim.x) {
            const double e = E[col + j];
            s += (prim[j] - e) * exp(-bt * (e + V[j] - V[m]));
        }
        s = wave_sum(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < nwave; ++w) t += red[w];
            prim[m] = t / m;
        }
        __syncthreads();
    }
    if (tid == 0) {
        double esum = 0.0;
        for (int m = 1; m <= N; ++m) esum += E[(size_t)(m - 1) * m / 2 + (m - 1)];   // E(m,1)
        out[0] = prim[N] / P;
        out[1] = exp(-bt * (esum - V[N]) - lgamma((double)N + 1.0));
        out[2] = exp(-bt * (E[(size_t)(N - 1) * N / 2] - V[N]));                     // E(N,N)
    }
}

}  // namespace pimdb
