// batch.hpp -- the batched persistent form: many independent ring polymers
// (uneven N, P, T, potentials' parameters) advanced by ONE kernel launch per
// MD step.  Each workgroup owns whole replicas, taken from a static schedule
// (longest-processing-time first, computed on the host), and runs the full
// OBABO step of a replica with workgroup barriers only:
//
//   O-step (+ zeroMomentum) -> B + A -> pair/ext/spring forces of all beads
//   -> cycle energies -> forward and backward recursions -> exchange forces
//   on beads 1 and P -> B -> O-step (+ zeroMomentum)
//
// (Simulation::run, src/simulation.cpp:246-259).  The device noise uses the
// same Philox keys as the single-system form, so a one-replica batch follows
// the single-system trajectory up to summation order.
#pragma once

#include "device_common.hpp"
#include "exchange.hpp"
#include "pairs.hpp"

namespace pimdb {

struct RepDesc {
    int N, P, bosonic, pbc;
    int thermostat, nm, fixcom, ext_harmonic;
    double L, invL, mass, dt;
    double k, bt, c1, c2;
    double ext_k, int_k, strength, cutoff;
    uint32_t seed, rid;
    unsigned long long off_x;   // P*D*N block of x, p, f
    unsigned long long off_E;   // N(N+1)/2 block of E, ET
    unsigned long long off_V;   // N+1 block of V, Vb
    unsigned long long off_C;   // P*P block of Ct, CTt
};

struct BatchPtrs {
    double *x, *p, *f, *E, *ET, *V, *Vb;
    const double *Ct, *CTt;
    const RepDesc* reps;
    const int* sched;          // replicas of workgroup g: sched[sched_off[g] .. sched_off[g+1])
    const int* sched_off;
    int* status;               // per replica: 1 when the exchange overflowed
    unsigned long long* ctl;   // [0] step counter, [1] finished-workgroup counter
};

// Sum over the workgroup of one double per thread; result in every thread.
__device__ __forceinline__ double wg_sum(double v, double* red) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

// Optional half kick, O-step (Cartesian or normal-mode coupling), optional
// zeroMomentum, for one replica (langevin.cpp:15-27, thermostat_coupling.cpp,
// simulation.cpp:581-603).
template <int D>
__device__ __noinline__ void wg_momenta(const RepDesc& r, const BatchPtrs& bp, bool kick, uint32_t step, uint32_t ostep,
                                        double* red, double* lds) {
    const int N = r.N, P = r.P;
    const int tot = P * D * N;
    double* p = bp.p + r.off_x;
    const double* f = bp.f + r.off_x;
    const double half_dt = 0.5 * r.dt;
    if (r.thermostat != 1 || !r.nm) {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            double v = p[e];
            if (kick) v += half_dt * f[e];
            if (r.thermostat == 1) {
                double g0, g1;
                gauss2((uint32_t)e, step, ostep, r.rid, r.seed, g0, g1);
                v = r.c1 * v + r.c2 * g0;
            }
            p[e] = v;
        }
    } else {
        if (kick) {
            for (int e = threadIdx.x; e < tot; e += blockDim.x) p[e] += half_dt * f[e];
            __syncthreads();
        }
        const double* Ct = bp.Ct + r.off_C;
        const double* CTt = bp.CTt + r.off_C;
        // Columns (a, i) in tiles of 64 staged in LDS; wave w computes modes
        // k = w, w + nwave, ... for the 64 columns of the tile (lane = column),
        // so C[k][j] is the same for the whole wave.
        const int ncols = D * N;
        const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
        double* pc = lds;                        // [P][64]
        double* pn = lds + (size_t)P * WAVE;     // [P][64]
        for (int c0 = 0; c0 < ncols; c0 += WAVE) {
            const int col = c0 + lane;
            const bool valid = col < ncols;
            for (int e = threadIdx.x; e < P * WAVE; e += blockDim.x) {
                const int j = e / WAVE, c = c0 + e % WAVE;
                pc[e] = c < ncols ? p[(size_t)j * ncols + c] : 0.0;
            }
            __syncthreads();
            // mode k of column (a, i): sum_j C[k][j] p[j][a][i], then the Langevin update
            for (int k = wave; k < P; k += nw) {
                double s = 0.0;
                for (int j = 0; j < P; ++j) s += Ct[(size_t)j * P + k] * pc[j * WAVE + lane];
                double g0 = 0.0, g1;
                if (valid) gauss2((uint32_t)((size_t)k * ncols + col), step, ostep, r.rid, r.seed, g0, g1);
                pn[k * WAVE + lane] = r.c1 * s + r.c2 * g0;
            }
            __syncthreads();
            for (int b = wave; b < P; b += nw) {
                double s = 0.0;
                for (int k = 0; k < P; ++k) s += CTt[(size_t)k * P + b] * pn[k * WAVE + lane];
                if (valid) p[(size_t)b * ncols + col] = s;
            }
            __syncthreads();
        }
    }
    if (r.fixcom) {
        __syncthreads();
        double cm[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double s = 0.0;
            for (int e = threadIdx.x; e < P * N; e += blockDim.x) {
                const int b = e / N, i = e % N;
                s += p[((size_t)b * D + a) * N + i];
            }
            cm[a] = wg_sum(s, red) / (double)(N * P);
        }
        for (int e = threadIdx.x; e < tot; e += blockDim.x) p[e] -= cm[(e / N) % D];
    }
    __syncthreads();
}

// B + A: p += dt/2 f, x += dt p / m (velocity_verlet.cpp:24-38)
template <int D>
__device__ __noinline__ void wg_kick_drift(const RepDesc& r, const BatchPtrs& bp) {
    const int tot = r.P * D * r.N;
    double* x = bp.x + r.off_x;
    double* p = bp.p + r.off_x;
    const double* f = bp.f + r.off_x;
    const double half_dt = 0.5 * r.dt;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const double pv = p[e] + half_dt * f[e];
        p[e] = pv;
        x[e] += r.dt * pv / r.mass;
    }
    __syncthreads();
}

// Physical forces and interior springs on every (bead, particle) of a replica
// (simulation.cpp:353-455).  Each wave owns whole beads: the pairs of a bead
// with Newton's third law (slice_pair_forces), then the trap and the springs.
template <int D, int KIND>
__device__ __noinline__ void wg_bead_forces(const RepDesc& r, const BatchPtrs& bp) {
    const int N = r.N, P = r.P;
    const double* x = bp.x + r.off_x;
    double* f = bp.f + r.off_x;
    const bool pbc = r.pbc != 0;
    const size_t bs = (size_t)D * N;
    const int lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const PairParams pp = make_pair_params(r.L, r.pbc, r.cutoff, r.int_k, r.strength);
    for (int b = wave; b < P; b += nw) {
        const double* xs = x + bs * b;
        double* fo = f + bs * b;
        if (KIND != POT_FREE && r.cutoff != 0.0) {
            slice_pair_forces<D, KIND>(xs, fo, N, pp, lane);
        } else {
            for (int e = lane; e < D * N; e += WAVE) fo[e] = 0.0;
            __threadfence_block();
        }
        // trap and interior springs; exchange beads get their spring part
        // from the exchange phase
        const bool exchange_bead = r.bosonic && (b == 0 || b == P - 1);
        const double* xprev = x + bs * ((b - 1 + P) % P);
        const double* xnext = x + bs * ((b + 1) % P);
        for (int e = lane; e < D * N; e += WAVE) {
            const double xi = xs[e];
            const double ext = r.ext_harmonic ? (xi * r.ext_k) * (-1.0) : -0.0;
            const double phys = ext + fo[e];
            double spring = 0.0;
            if (!exchange_bead) {
                const double dp = disp(xi, xprev[e], pbc, r.L);
                const double dn = disp(xi, xnext[e], pbc, r.L);
                spring = r.k * (dp + dn);
            }
            fo[e] = exchange_bead ? phys : spring + phys;
        }
    }
    __syncthreads();
}

// LDS doubles of the integrator's normal-mode column tiles.
__host__ __device__ inline size_t batch_nm_lds_doubles(int P) { return 2 * (size_t)P * WAVE; }

// Exchange of one replica: cycle energies, both recursions, exterior forces
// added into f of beads 1 and P.
template <int D>
__device__ __noinline__ void wg_exchange(const RepDesc& r, const BatchPtrs& bp, double* lds, int ri) {
    const int N = r.N, P = r.P;
    const double* x = bp.x + r.off_x;
    double* f = bp.f + r.off_x;
    double* E = bp.E + r.off_E;
    double* ET = bp.ET + r.off_E;
    double* V = bp.V + r.off_V;
    double* Vb = bp.Vb + r.off_V;
    const size_t bs = (size_t)D * N;
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    const double* x1 = x;
    const double* xP = x + bs * (P - 1);
    for (int v = wave; v < N; v += nw) cycle_energy_column<D>(v, N, lane, x1, xP, r.k, r.pbc != 0, r.L, E, ET);
    __threadfence_block();
    __syncthreads();
    exchange_recursion<0>(N, r.bt, E, ET, V, Vb, bp.status + ri, rec_carve(lds, N));
    exchange_recursion<1>(N, r.bt, E, ET, V, Vb, bp.status + ri, rec_carve(lds, N));
    __threadfence_block();
    const double* x2 = x + bs * (1 % P);
    const double* xPm1 = x + bs * ((P - 2 + P) % P);
    for (int task = wave; task < 2 * N; task += nw) {
        const int which = task / N, l = task % N;
        double F[D];
        exterior_force<D>(which, l, N, lane, r.bt, r.k, r.pbc != 0, r.L, E, ET, V, Vb, x1, x2, xPm1, xP, false, F);
        if (lane == 0) {
            const size_t b = which == 0 ? 0 : (size_t)(P - 1);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const size_t idx = (b * D + a) * N + l;
                f[idx] = F[a] + f[idx];
            }
        }
    }
    __syncthreads();
}

// Work of one launch: a full MD step, or parts of it for timing.
enum BatchMode : int { BM_STEP = 0, BM_FORCES = 1, BM_PAIRS = 2, BM_EXCHANGE = 3, BM_INTEGRATOR = 4 };

// One MD step of every replica (BM_STEP), the forces only (BM_FORCES), or one
// phase for timing.  MAXT is the largest workgroup size the instantiation is
// compiled for (it sets the register budget).
template <int D, int KIND, int MAXT>
static __global__ void __launch_bounds__(MAXT) k_batch_step(BatchPtrs bp, int mode) {
    extern __shared__ double smem[];
    double* red = smem;                 // [16]
    double* lds = smem + 16;            // normal-mode tiles / recursion workspace (not live together)
    const uint32_t step = (uint32_t)bp.ctl[0];
    const int g = blockIdx.x;
    for (int s = bp.sched_off[g]; s < bp.sched_off[g + 1]; ++s) {
        const int ri = bp.sched[s];
        const RepDesc& r = bp.reps[ri];
        const bool integ = mode == BM_STEP || mode == BM_INTEGRATOR;
        if (integ) {
            wg_momenta<D>(r, bp, false, step, 0, red, lds);
            wg_kick_drift<D>(r, bp);
        }
        if (mode <= BM_PAIRS) wg_bead_forces<D, KIND>(r, bp);
        if (r.bosonic && (mode <= BM_FORCES || mode == BM_EXCHANGE)) wg_exchange<D>(r, bp, lds, ri);
        if (integ) wg_momenta<D>(r, bp, true, step + 1, 1, red, lds);
    }
    if (mode == BM_STEP) {
        // the last workgroup to finish advances the step counter
        __shared__ int last;
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const unsigned long long done = atomicAdd(&bp.ctl[1], 1ull);
            last = (done == (unsigned long long)gridDim.x - 1);
        }
        __syncthreads();
        if (last && threadIdx.x == 0) {
            atomicExch(&bp.ctl[1], 0ull);
            atomicAdd(&bp.ctl[0], 1ull);
        }
    }
}

}  // namespace pimdb
