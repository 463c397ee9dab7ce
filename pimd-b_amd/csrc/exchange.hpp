// exchange.hpp -- Feldman-Hirshberg bosonic exchange on the GPU: cycle
// energies, the forward/backward recursions and the exchange forces (the
// adjoint of V_B plus the interior links) on beads 1 and P.
//
// Reference: src/bosonic_exchange/quadratic_bosonic_exchange.cpp and
// bosonic_exchange_base.cpp of higj/pimd-b.
//
// Storage: bead-major SoA positions, x[(b*D + a)*N + i].  The cycle energy of
// the cycle through particles u..v (u <= v) is ecyc(u,v) = E^(v+1, v-u+1) of
// the reference and lives at E[v(v+1)/2 + u]: this is exactly the reference's
// packed E_kn triangle (E(m,k) at m(m+1)/2 - k), so column v is contiguous.
//
// Recursions.  Both recursions have the form
//     U_n = -ln sum_{q<n} exp(-(w(n,q) + U_q)),  U_0 = 0,
// forward : U_n = beta~ V[n],      w(n,q) = beta~ ecyc(q, n-1) + ln n
// backward: U_n = beta~ Vb[N-n],   w(n,q) = beta~ ecyc(N-n, N-1-q) + ln(N-q)
// (quadratic_bosonic_exchange.cpp:73-128).  Instead of one log-sum-exp per n
// (a reduction, an exp and a log on the sequential path), targets are taken in
// blocks of REC_B and each block runs
//   (1) a min-plus pass  R_n = min_q (w(n,q) + Uhat_q)   [sequential part: add+min]
//   (2) a linear solve   z_n = sum_q exp(-(w(n,q) + Uhat_q - R_n)) zhat_q
//                                                         [sequential part: one FMA]
//   (3) U_n = R_n - ln z_n
// where Uhat_q = U_q, zhat_q = 1 for sources of earlier blocks and
// Uhat_q = R_q, zhat_q = z_q inside the block.  Every exponent is <= 0 and the
// minimising term contributes exactly zhat >= 1, so 1 <= z_n <= n 2^REC_B: no
// overflow or underflow for any N.  All exponentials are off the sequential
// path.  The result equals the reference's shifted log-sum-exp up to rounding.
#pragma once

#include "device_common.hpp"

namespace pimdb {

constexpr int REC_B = 32;   // targets per block (one wave does the sequential parts)

// Cycle energies of column v, one wave per column.  ecyc(u,v) =
// k/2 (sum_{w=u}^{v-1} |x1[w+1]-xP[w]|^2 + |xP[v]-x1[u]|^2): the closed form of
// the reference's recurrence (quadratic_bosonic_exchange.cpp:45-60, whose
// telescoping terms cancel), summed by a wave scan from u = v downwards so the
// rounding error is relative to each entry, not to the whole column.
template <int D>
__device__ __forceinline__ void cycle_energy_column(int v, int N, int lane, const double* __restrict__ x1,
                                                    const double* __restrict__ xP, double k, bool pbc,
                                                    double L, double* __restrict__ E) {
    double xpv[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xpv[a] = xP[a * N + v];
    double carry = 0.0;
    const size_t col = (size_t)v * (v + 1) / 2;
    for (int base = 0; base <= v; base += WAVE) {
        const int u = v - (base + lane);
        double t = 0.0, c = 0.0;
        if (u >= 0) {
            double d[D];
#pragma unroll
            for (int a = 0; a < D; ++a) d[a] = disp(x1[a * N + u], xpv[a], pbc, L);
            c = norm2<D>(d);
            if (u < v) {
                double e[D];
#pragma unroll
                for (int a = 0; a < D; ++a) e[a] = disp(xP[a * N + u], x1[a * N + u + 1], pbc, L);
                t = norm2<D>(e);
            }
        }
        // inclusive scan over lanes (u descending)
        double s = t;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const double o = __shfl_up(s, off, WAVE);
            if (lane >= off) s += o;
        }
        if (u >= 0) E[col + u] = 0.5 * k * ((carry + s) + c);
        carry += __shfl(s, WAVE - 1, WAVE);
    }
}

// x1, xP: beads 1 and P, SoA [a][i].
template <int D>
static __global__ void __launch_bounds__(256) k_cycle_energies(int N, const double* __restrict__ x1,
                                                        const double* __restrict__ xP, double k, int pbc, double L,
                                                        double* __restrict__ E) {
    const int v = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (v >= N) return;
    cycle_energy_column<D>(v, N, threadIdx.x % WAVE, x1, xP, k, pbc != 0, L, E);
}

// Shared-memory workspace of one workgroup running both recursions.
struct RecShared {
    double* U[2];      // [N+1] each: U of forward / backward
    double* lg;        // [N+1]  ln(i)
    double* R[2];      // [REC_B]
    double* z[2];      // [REC_B]
    double* acc[2];    // [REC_B]  phase partials
    double* a[2];      // [REC_B*REC_B] in-block coefficients a[j*REC_B + t]
};

__device__ __forceinline__ size_t rec_shared_doubles(int N) {
    return 3 * (size_t)(N + 1) + 6 * REC_B + 2 * (size_t)REC_B * REC_B;
}

__device__ __forceinline__ RecShared rec_carve(double* base, int N) {
    RecShared s;
    double* p = base;
    s.U[0] = p; p += N + 1;
    s.U[1] = p; p += N + 1;
    s.lg = p; p += N + 1;
    for (int d = 0; d < 2; ++d) { s.R[d] = p; p += REC_B; }
    for (int d = 0; d < 2; ++d) { s.z[d] = p; p += REC_B; }
    for (int d = 0; d < 2; ++d) { s.acc[d] = p; p += REC_B; }
    for (int d = 0; d < 2; ++d) { s.a[d] = p; p += (size_t)REC_B * REC_B; }
    return s;
}

// w(n,q) of direction dir (see file header).
__device__ __forceinline__ double rec_w(int dir, int N, const double* __restrict__ E, double bt,
                                        const double* lg, int n, int q) {
    if (dir == 0) return bt * E[(size_t)(n - 1) * n / 2 + q] + lg[n];
    const int p = N - 1 - q;
    return bt * E[(size_t)p * (p + 1) / 2 + (N - n)] + lg[N - q];
}

// Both recursions for one ring, executed by the whole workgroup (any
// multiple of 128 threads).  Writes V[0..N], Vb[0..N] and sets *status to 1
// when a potential is not finite (the reference throws overflow_error,
// quadratic_bosonic_exchange.cpp:92-97, 119-124).
__device__ void exchange_recursions(int N, double bt, const double* __restrict__ E, double* __restrict__ V,
                                    double* __restrict__ Vb, int* __restrict__ status, RecShared sh) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid % WAVE, wave = tid / WAVE, nwave = nthr / WAVE;
    const int ntarget[2] = {N, N - 1};
    for (int i = tid; i <= N; i += nthr) {
        sh.lg[i] = (i > 0) ? log((double)i) : 0.0;
        sh.U[0][i] = 0.0;
        sh.U[1][i] = 0.0;
    }
    __syncthreads();
    for (int n0 = 1; n0 <= N; n0 += REC_B) {
        int nb[2];
        for (int d = 0; d < 2; ++d) nb[d] = max(0, min(REC_B, ntarget[d] - n0 + 1));
        // (0) in-block weights w(n0+t, n0+j), j < t, staged in LDS as a[j][t]
        for (int idx = tid; idx < 2 * REC_B * REC_B; idx += nthr) {
            const int d = idx / (REC_B * REC_B);
            const int j = (idx / REC_B) % REC_B, t = idx % REC_B;
            sh.a[d][j * REC_B + t] = (t < nb[d] && j < t) ? rec_w(d, N, E, bt, sh.lg, n0 + t, n0 + j) : 0.0;
        }
        // (1a) min over sources of earlier blocks, one wave per (dir, target)
        for (int task = wave; task < 2 * REC_B; task += nwave) {
            const int d = task / REC_B, t = task % REC_B;
            if (t >= nb[d]) continue;
            const int n = n0 + t;
            double m = INFINITY;
            for (int q = lane; q < n0; q += WAVE) m = fmin(m, rec_w(d, N, E, bt, sh.lg, n, q) + sh.U[d][q]);
            m = wave_min(m);
            if (lane == 0) sh.acc[d][t] = m;
        }
        __syncthreads();
        // (1b) in-block min-plus: wave 0 forward, wave 1 backward
        if (wave < 2) {
            const int d = wave;
            const int t = lane;
            const double* wb = sh.a[d];
            double r = (t < nb[d]) ? sh.acc[d][t] : INFINITY;
            double wnext = wb[t];
            for (int j = 0; j + 1 < nb[d]; ++j) {
                const double wj = wnext;
                wnext = wb[(j + 1) * REC_B + t];
                const double Rj = wave_bcast_uniform(r, j);
                if (t > j) r = fmin(r, wj + Rj);
            }
            if (t < nb[d]) sh.R[d][t] = r;
        }
        __syncthreads();
        // (2a) sum over sources of earlier blocks
        for (int task = wave; task < 2 * REC_B; task += nwave) {
            const int d = task / REC_B, t = task % REC_B;
            if (t >= nb[d]) continue;
            const int n = n0 + t;
            const double Rn = sh.R[d][t];
            double s = 0.0;
            for (int q = lane; q < n0; q += WAVE) s += exp(-(rec_w(d, N, E, bt, sh.lg, n, q) + sh.U[d][q] - Rn));
            s = wave_sum(s);
            if (lane == 0) sh.z[d][t] = s;
        }
        // (2b) in-block coefficients a[j][t] = exp(-(w(n0+t, n0+j) + R_j - R_t)), j < t
        for (int idx = tid; idx < 2 * REC_B * REC_B; idx += nthr) {
            const int d = idx / (REC_B * REC_B);
            const int j = (idx / REC_B) % REC_B, t = idx % REC_B;
            if (t < nb[d] && j < t) {
                double* ap = &sh.a[d][j * REC_B + t];
                *ap = exp(-(*ap + sh.R[d][j] - sh.R[d][t]));
            }
        }
        __syncthreads();
        // (2c) in-block forward substitution
        if (wave < 2) {
            const int d = wave;
            const int t = lane;
            const double* ab = sh.a[d];
            double z = (t < nb[d]) ? sh.z[d][t] : 0.0;
            double anext = ab[t];
            for (int j = 0; j + 1 < nb[d]; ++j) {
                const double aj = anext;
                anext = ab[(j + 1) * REC_B + t];
                const double zj = wave_bcast_uniform(z, j);
                z = fma(aj, zj, z);
            }
            if (t < nb[d]) {
                const double u = sh.R[d][t] - log(z);
                sh.U[d][n0 + t] = u;
                if (!isfinite(u)) atomicOr(status, 1);
            }
        }
        __syncthreads();
    }
    // outputs in the reference's units
    for (int i = tid; i <= N; i += nthr) {
        V[i] = sh.U[0][i] / bt;
        if (i >= 1 && i < N) Vb[i] = sh.U[1][N - i] / bt;
    }
    if (tid == 0) {
        Vb[N] = 0.0;
        Vb[0] = sh.U[0][N] / bt;   // V_backwards[0] = V[N], :127
    }
    __syncthreads();
}

static __global__ void __launch_bounds__(1024) k_exchange_recursions(int N, double bt, const double* __restrict__ E,
                                                              double* __restrict__ V, double* __restrict__ Vb,
                                                              int* __restrict__ status) {
    extern __shared__ double smem[];
    exchange_recursions(N, bt, E, V, Vb, status, rec_carve(smem, N));
}

// Connection probability P[l][u] (quadratic_bosonic_exchange.cpp:142-157).
__device__ __forceinline__ double conn_prob(int N, double bt, const double* __restrict__ E,
                                            const double* __restrict__ V, const double* __restrict__ Vb,
                                            int l, int u) {
    if (u == l + 1 && l < N - 1) return 1.0 - (exp(-bt * (V[l + 1] + Vb[l + 1] - V[N])));
    if (u <= l) {
        const double e = E[(size_t)l * (l + 1) / 2 + u];
        return 1.0 / (l + 1) * exp(-bt * (V[u] + e + Vb[l + 1] - V[N]));
    }
    return 0.0;
}

static __global__ void __launch_bounds__(256) k_exchange_probabilities(int N, double bt, const double* __restrict__ E,
                                                                const double* __restrict__ V,
                                                                const double* __restrict__ Vb,
                                                                double* __restrict__ prob) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)N * N) return;
    const int l = (int)(idx / N), u = (int)(idx % N);
    prob[idx] = conn_prob(N, bt, E, V, Vb, l, u);
}

// Exterior spring force of particle l on bead 1 (which=0) or bead P
// (which=1), one wave per (which, l):
//  bead 1: F[l] = k [ sum_{j>=max(0,l-1)} P[j][l] (xP[j]-x1[l]) + (x2[l]-x1[l]) ]   (:188-215)
//  bead P: F[l] = k [ sum_{j<=min(l+1,N-1)} P[l][j] (x1[j]-xP[l]) + (xPm1[l]-xP[l]) ] (:159-186)
// Probabilities are recomputed from V, Vb and E.  With grad_only the
// interior link is left out and the result is -dV_B/dx (the adjoint of V_B
// alone, used by the autograd harness).
template <int D>
__device__ __forceinline__ void exterior_force(int which, int l, int N, int lane, double bt, double k, bool pbc,
                                               double L, const double* __restrict__ E, const double* __restrict__ V,
                                               const double* __restrict__ Vb, const double* __restrict__ x1,
                                               const double* __restrict__ x2, const double* __restrict__ xPm1,
                                               const double* __restrict__ xP, bool grad_only, double (&F)[D]) {
    const double* self = which == 0 ? x1 : xP;
    const double* other = which == 0 ? xP : x1;
    double xs[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xs[a] = self[a * N + l];
    double S[D];
#pragma unroll
    for (int a = 0; a < D; ++a) S[a] = 0.0;
    const int jlo = which == 0 ? max(0, l - 1) : 0;
    const int jhi = which == 0 ? N - 1 : min(l + 1, N - 1);
    for (int j = jlo + lane; j <= jhi; j += WAVE) {
        const double pr = which == 0 ? conn_prob(N, bt, E, V, Vb, j, l) : conn_prob(N, bt, E, V, Vb, l, j);
#pragma unroll
        for (int a = 0; a < D; ++a) S[a] += pr * disp(xs[a], other[a * N + j], pbc, L);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) S[a] = wave_sum(S[a]);
    const double* link = which == 0 ? x2 : xPm1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double s = S[a];
        if (!grad_only) s += disp(xs[a], link[a * N + l], pbc, L);
        F[a] = s * k;
    }
}

// Adds the exterior spring forces into f for beads 1 and P:
// f = F_exchange + f_phys (simulation.cpp:363-373: spring + physical).
template <int D>
static __global__ void __launch_bounds__(256) k_exchange_forces(int N, int P, double bt, double k, int pbc, double L,
                                                         const double* __restrict__ E, const double* __restrict__ V,
                                                         const double* __restrict__ Vb, const double* __restrict__ x,
                                                         double* __restrict__ f, double* __restrict__ fex) {
    const int task = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (task >= 2 * N) return;
    const int which = task / N, l = task % N;
    const int lane = threadIdx.x % WAVE;
    const size_t bs = (size_t)D * N;
    const double* x1 = x;
    const double* x2 = x + bs * (1 % P);
    const double* xPm1 = x + bs * ((P - 2 + P) % P);
    const double* xP = x + bs * (P - 1);
    double F[D];
    exterior_force<D>(which, l, N, lane, bt, k, pbc != 0, L, E, V, Vb, x1, x2, xPm1, xP, false, F);
    if (lane == 0) {
        const size_t b = which == 0 ? 0 : (size_t)(P - 1);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const size_t idx = (b * D + a) * N + l;
            if (fex) fex[((size_t)which * D + a) * N + l] = F[a];
            if (f) f[idx] = F[a] + f[idx];
        }
    }
}

// -dV_B/dx1 and -dV_B/dxP on two N x D slices (autograd harness): x1, xP
// given separately in SoA [a][i]; out[which][a][i].
template <int D>
static __global__ void __launch_bounds__(256) k_exchange_vb_grad(int N, double bt, double k, int pbc, double L,
                                                          const double* __restrict__ E, const double* __restrict__ V,
                                                          const double* __restrict__ Vb, const double* __restrict__ x1,
                                                          const double* __restrict__ xP, double* __restrict__ out) {
    const int task = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (task >= 2 * N) return;
    const int which = task / N, l = task % N;
    const int lane = threadIdx.x % WAVE;
    double F[D];
    exterior_force<D>(which, l, N, lane, bt, k, pbc != 0, L, E, V, Vb, x1, x1, xP, xP, true, F);
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) out[((size_t)which * D + a) * N + l] = F[a];
    }
}

}  // namespace pimdb
