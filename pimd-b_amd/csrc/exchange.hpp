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

constexpr int REC_B = 64;   // targets per block: one per lane of the sequential wave
constexpr int REC_MAXWAVES = 16;
constexpr int REC_UNROLL = 8;  // independent source loads in flight per lane

// Row u of the row-major copy ET starts at u N - u(u-1)/2 and holds v = u..N-1.
__host__ __device__ __forceinline__ size_t et_row(int u, int N) { return (size_t)u * N - (size_t)u * (u - 1) / 2; }

// Cycle energies of column v, one wave per column.  ecyc(u,v) =
// k/2 (sum_{w=u}^{v-1} |x1[w+1]-xP[w]|^2 + |xP[v]-x1[u]|^2): the closed form of
// the reference's recurrence (quadratic_bosonic_exchange.cpp:45-60, whose
// telescoping terms cancel), summed by a wave scan from u = v downwards so the
// rounding error is relative to each entry, not to the whole column.
template <int D>
__device__ __forceinline__ void cycle_energy_column(int v, int N, int lane, const double* __restrict__ x1,
                                                    const double* __restrict__ xP, double k, bool pbc,
                                                    double L, double* __restrict__ E, double* __restrict__ ET) {
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
        if (u >= 0) {
            const double e = 0.5 * k * ((carry + s) + c);
            E[col + u] = e;
            if (ET) ET[et_row(u, N) + (v - u)] = e;
        }
        carry += __shfl(s, WAVE - 1, WAVE);
    }
}

// x1, xP: beads 1 and P, SoA [a][i].  ET may be null.
template <int D>
static __global__ void __launch_bounds__(256) k_cycle_energies(int N, const double* __restrict__ x1,
                                                        const double* __restrict__ xP, double k, int pbc, double L,
                                                        double* __restrict__ E, double* __restrict__ ET) {
    const int v = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (v >= N) return;
    cycle_energy_column<D>(v, N, threadIdx.x % WAVE, x1, xP, k, pbc != 0, L, E, ET);
}

// Shared-memory workspace of one workgroup running one recursion.
struct RecShared {
    double* U;         // [N+1]
    double* lg;        // [N+1]  ln(i)
    double* R;         // [REC_B] shifts of the block's targets
    double* part;      // [REC_MAXWAVES][REC_B] per-wave partial min / sum
    double* a;         // [REC_B*REC_B] in-block weights, then coefficients, a[j*REC_B + t]
};

__host__ __device__ __forceinline__ size_t rec_shared_doubles(int N) {
    return 2 * (size_t)(N + 1) + REC_B + (size_t)REC_MAXWAVES * REC_B + (size_t)REC_B * REC_B;
}

__device__ __forceinline__ RecShared rec_carve(double* base, int N) {
    RecShared s;
    double* p = base;
    s.U = p; p += N + 1;
    s.lg = p; p += N + 1;
    s.R = p; p += REC_B;
    s.part = p; p += (size_t)REC_MAXWAVES * REC_B;
    s.a = p;
    return s;
}

// w(n,q) of direction dir (see file header).  The forward direction reads
// the row-major copy ET (row q, contiguous in n), the backward one reads E
// (column N-1-q, contiguous in n): consecutive targets hit consecutive
// addresses in both.
template <int DIR>
__device__ __forceinline__ double rec_w(int N, const double* __restrict__ E, const double* __restrict__ ET, double bt,
                                        const double* lg, int n, int q) {
    if constexpr (DIR == 0) {
        return bt * ET[et_row(q, N) + (n - 1 - q)] + lg[n];
    } else {
        const int p = N - 1 - q;
        return bt * E[(size_t)p * (p + 1) / 2 + (N - n)] + lg[N - q];
    }
}

// One recursion (DIR 0: forward V, DIR 1: backward Vb) for one ring, run by
// the whole workgroup (64..1024 threads).  Lane t of every wave serves target
// n0+t of the current block; the waves split the sources of earlier blocks
// and keep UNROLL loads in flight.  Writes V[0..N] and Vb[0] (DIR 0) or
// Vb[1..N] (DIR 1); sets *status to 1 when a potential is not finite (the
// reference throws overflow_error, quadratic_bosonic_exchange.cpp:92-97,
// 119-124).
template <int DIR, int UNROLL = REC_UNROLL>
__device__ void exchange_recursion(int N, double bt, const double* __restrict__ E, const double* __restrict__ ET,
                                   double* __restrict__ V, double* __restrict__ Vb, int* __restrict__ status,
                                   RecShared sh) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid % WAVE, wave = tid / WAVE;
    const int nwave = min(nthr / WAVE, REC_MAXWAVES);
    const int ntarget = DIR == 0 ? N : N - 1;
    for (int i = tid; i <= N; i += nthr) {
        sh.lg[i] = (i > 0) ? log((double)i) : 0.0;
        sh.U[i] = 0.0;
    }
    __syncthreads();
    for (int n0 = 1; n0 <= ntarget; n0 += REC_B) {
        const int nb = min(REC_B, ntarget - n0 + 1);
        // (0) in-block weights w(n0+t, n0+j), j < t, staged as a[j][t]
        for (int idx = tid; idx < REC_B * REC_B; idx += nthr) {
            const int j = idx / REC_B, t = idx % REC_B;
            sh.a[idx] = (t < nb && j < t) ? rec_w<DIR>(N, E, ET, bt, sh.lg, n0 + t, n0 + j) : 0.0;
        }
        // (1a) per-wave partial min over this wave's share of earlier sources
        const int chunk = (n0 + nwave - 1) / nwave;
        const int qlo = min(n0, wave * chunk), qhi = min(n0, qlo + chunk);
        const bool active = wave < nwave && lane < nb;
        if (wave < nwave) {
            double m = INFINITY;
            if (active) {
                int q = qlo;
                for (; q + UNROLL <= qhi; q += UNROLL) {
                    double w[UNROLL];
#pragma unroll
                    for (int r = 0; r < UNROLL; ++r) w[r] = rec_w<DIR>(N, E, ET, bt, sh.lg, n0 + lane, q + r);
#pragma unroll
                    for (int r = 0; r < UNROLL; ++r) m = fmin(m, w[r] + sh.U[q + r]);
                }
                for (; q < qhi; ++q) m = fmin(m, rec_w<DIR>(N, E, ET, bt, sh.lg, n0 + lane, q) + sh.U[q]);
            }
            sh.part[wave * REC_B + lane] = m;
        }
        __syncthreads();
        // (1b) in-block min-plus on wave 0
        if (wave == 0) {
            const int t = lane;
            double r = INFINITY;
            for (int w = 0; w < nwave; ++w) r = fmin(r, sh.part[w * REC_B + t]);
            if (t >= nb) r = INFINITY;
            double wnext = sh.a[t];
            for (int j = 0; j + 1 < nb; ++j) {
                const double wj = wnext;
                wnext = sh.a[(j + 1) * REC_B + t];
                const double Rj = wave_bcast_uniform(r, j);
                if (t > j) r = fmin(r, wj + Rj);
            }
            sh.R[t] = r;
        }
        __syncthreads();
        // (2a) per-wave partial sums over earlier sources
        if (wave < nwave) {
            double s = 0.0;
            if (active) {
                const double Rn = sh.R[lane];
                int q = qlo;
                for (; q + UNROLL <= qhi; q += UNROLL) {
                    double w[UNROLL];
#pragma unroll
                    for (int r = 0; r < UNROLL; ++r) w[r] = rec_w<DIR>(N, E, ET, bt, sh.lg, n0 + lane, q + r);
#pragma unroll
                    for (int r = 0; r < UNROLL; ++r) s += exp(-(w[r] + sh.U[q + r] - Rn));
                }
                for (; q < qhi; ++q) s += exp(-(rec_w<DIR>(N, E, ET, bt, sh.lg, n0 + lane, q) + sh.U[q] - Rn));
            }
            sh.part[wave * REC_B + lane] = s;
        }
        // (2b) in-block coefficients a[j][t] = exp(-(w(n0+t, n0+j) + R_j - R_t)), j < t
        for (int idx = tid; idx < REC_B * REC_B; idx += nthr) {
            const int j = idx / REC_B, t = idx % REC_B;
            if (t < nb && j < t) sh.a[idx] = exp(-(sh.a[idx] + sh.R[j] - sh.R[t]));
        }
        __syncthreads();
        // (2c) in-block forward substitution on wave 0, then U = R - ln z
        if (wave == 0) {
            const int t = lane;
            double z = 0.0;
            for (int w = 0; w < nwave; ++w) z += sh.part[w * REC_B + t];
            double anext = sh.a[t];
            for (int j = 0; j + 1 < nb; ++j) {
                const double aj = anext;
                anext = sh.a[(j + 1) * REC_B + t];
                const double zj = wave_bcast_uniform(z, j);
                z = fma(aj, zj, z);
            }
            if (t < nb) {
                const double u = sh.R[t] - log(z);
                sh.U[n0 + t] = u;
                if (!isfinite(u)) atomicOr(status, 1);
            }
        }
        __syncthreads();
    }
    // outputs in the reference's units
    if constexpr (DIR == 0) {
        for (int i = tid; i <= N; i += nthr) V[i] = sh.U[i] / bt;
        if (tid == 0) Vb[0] = sh.U[N] / bt;   // V_backwards[0] = V[N], :127
    } else {
        for (int i = tid; i <= N; i += nthr) {
            if (i >= 1 && i < N) Vb[i] = sh.U[N - i] / bt;
        }
        if (tid == 0) Vb[N] = 0.0;
    }
    __syncthreads();
}

// Block 0 runs the forward recursion, block 1 the backward one.
static __global__ void __launch_bounds__(1024) k_exchange_recursions(int N, double bt, const double* __restrict__ E,
                                                                     const double* __restrict__ ET,
                                                                     double* __restrict__ V, double* __restrict__ Vb,
                                                                     int* __restrict__ status) {
    extern __shared__ double smem[];
    if (blockIdx.x == 0) exchange_recursion<0>(N, bt, E, ET, V, Vb, status, rec_carve(smem, N));
    else exchange_recursion<1>(N, bt, E, ET, V, Vb, status, rec_carve(smem, N));
}

// Connection probability P[l][u] (quadratic_bosonic_exchange.cpp:142-157).
// With ET given, ecyc(u,l) is read from the row-major copy.
__device__ __forceinline__ double conn_prob(int N, double bt, const double* __restrict__ E,
                                            const double* __restrict__ ET, const double* __restrict__ V,
                                            const double* __restrict__ Vb, int l, int u) {
    if (u == l + 1 && l < N - 1) return 1.0 - (exp(-bt * (V[l + 1] + Vb[l + 1] - V[N])));
    if (u <= l) {
        const double e = ET ? ET[et_row(u, N) + (l - u)] : E[(size_t)l * (l + 1) / 2 + u];
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
    prob[idx] = conn_prob(N, bt, E, nullptr, V, Vb, l, u);
}

// Exterior spring force of particle l on bead 1 (which=0) or bead P
// (which=1), one wave per (which, l):
//  bead 1: F[l] = k [ sum_{j>=max(0,l-1)} P[j][l] (xP[j]-x1[l]) + (x2[l]-x1[l]) ]   (:188-215)
//  bead P: F[l] = k [ sum_{j<=min(l+1,N-1)} P[l][j] (x1[j]-xP[l]) + (xPm1[l]-xP[l]) ] (:159-186)
// Probabilities are recomputed from V, Vb and E (ET required).  With grad_only the
// interior link is left out and the result is -dV_B/dx (the adjoint of V_B
// alone, used by the autograd harness).
template <int D, int JUNROLL = 4>
__device__ __forceinline__ void exterior_force(int which, int l, int N, int lane, double bt, double k, bool pbc,
                                               double L, const double* __restrict__ E, const double* __restrict__ ET,
                                               const double* __restrict__ V,
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
    // P[row][col] (conn_prob's expression) with (row, col) = (j, l) for bead
    // 1 and (l, j) for bead P; the cycle energy ecyc(col, row) it needs is
    // ebase[j] in both cases (row l of ET, or row l of the packed E).
    const double* ebase = which == 0 ? ET + et_row(l, N) - l : E + (size_t)l * (l + 1) / 2;
    const double VN = V[N];
#pragma unroll JUNROLL
    for (int j = jlo + lane; j <= jhi; j += WAVE) {
        const int row = which == 0 ? j : l, col = which == 0 ? l : j;
        const bool link = col == row + 1 && row < N - 1;
        double pr = 0.0;
        if (link || col <= row) {
            const double arg = link ? V[row + 1] + Vb[row + 1] - VN : V[col] + ebase[j] + Vb[row + 1] - VN;
            const double ex = exp(-bt * arg);
            pr = link ? 1.0 - ex : 1.0 / (row + 1) * ex;
        }
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
                                                         const double* __restrict__ E, const double* __restrict__ ET,
                                                         const double* __restrict__ V,
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
    exterior_force<D>(which, l, N, lane, bt, k, pbc != 0, L, E, ET, V, Vb, x1, x2, xPm1, xP, false, F);
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
                                                          const double* __restrict__ E, const double* __restrict__ ET,
                                                          const double* __restrict__ V,
                                                          const double* __restrict__ Vb, const double* __restrict__ x1,
                                                          const double* __restrict__ xP, double* __restrict__ out) {
    const int task = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (task >= 2 * N) return;
    const int which = task / N, l = task % N;
    const int lane = threadIdx.x % WAVE;
    double F[D];
    exterior_force<D>(which, l, N, lane, bt, k, pbc != 0, L, E, ET, V, Vb, x1, x1, xP, xP, true, F);
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) out[((size_t)which * D + a) * N + l] = F[a];
    }
}

}  // namespace pimdb
