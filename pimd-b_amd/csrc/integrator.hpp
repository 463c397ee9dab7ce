// integrator.hpp -- OBABO velocity Verlet, Langevin O-steps (Cartesian or
// normal-mode coupling), the normal-mode propagator and centre-of-mass
// momentum removal, on bead-major SoA arrays [b][a][i].
//
// Reference: Simulation::run (src/simulation.cpp:246-259), zeroMomentum
// (:581-603), Propagator::momentStep/coordsStep and VelocityVerletPropagator
// (src/propagators/velocity_verlet.cpp:7-38), NormalModesPropagator
// (src/propagators/normal_modes_propagator.cpp:19-103), NormalModes
// (src/normal_modes.cpp), LangevinThermostat (src/thermostats/langevin.cpp)
// and the couplings (src/thermostats/thermostat_coupling.cpp).
//
// Momentum columns (all P beads of one (axis, particle)) are the unit of the
// normal-mode transforms; a workgroup owns a tile of TC particles of one axis
// for all beads, stages it in LDS and applies the P x P matrices there.  The
// transforms are small (P <= 256) dense products kept on the FP64 VALU.
#pragma once

#include "device_common.hpp"

namespace pimdb {

struct MomentaArgs {
    int N, P, D, tc;            // tc: particles per tile
    double half_dt;             // 0.5 * dt (momentStep, velocity_verlet.cpp:29)
    double c1, c2;              // Langevin friction / noise (langevin.cpp:10-13)
    int thermostat;             // 0 none, 1 langevin
    int nm;                     // normal-mode coupling (thermostat_coupling.cpp:19-47)
    const double* Ct;           // cart->nm matrix transposed: Ct[j*P + k] = C[k][j]
    const double* CTt;          // nm->cart matrix transposed: CTt[k*P + b] = CT[b][k]
    const double* noise;        // host noise [b][a][i] for this O-step, or null (Philox)
    const unsigned long long* step_ctr;
    uint32_t ostep, seed, replica;
    const double* com_in;       // partial momentum sums to subtract first, or null
    int ncom_in;                // number of partial rows [ncom_in][D]
    double* com_out;            // partial sums of the result, [gridDim.x][D], or null
    const double* kick_f;       // forces for a half kick first, or null
    // 0: the whole O-step.  Normal-mode coupling with a thermostat that acts
    // on all modes at once (Nose-Hoover): 1 writes the modes of the staged
    // momenta to `modes` [k][a][i] and stops; 2 reads the (thermostatted)
    // modes back, transforms them to p and writes the partial sums.
    int stage;
    double* modes;
};

// Centre-of-mass momentum per axis from the partial sums [npart][D], reduced
// once per workgroup into scm[0..D-1] (wave a reduces axis a); every
// workgroup runs the same reduction tree, so all see the same values.  Must
// be called by all threads of the workgroup.
__device__ __forceinline__ void block_com(const double* part, int npart, int D, int N, int P, double* scm) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (wave < D) {
        double s = 0.0;
        for (int t = lane; t < npart; t += WAVE) s += part[(size_t)t * D + wave];
        s = wave_sum(s);
        if (lane == 0) scm[wave] = s / (double)(N * P);
    }
    __syncthreads();
}

// Langevin noise of elements (2m, a, i) and (2m+1, a, i) (beads, or modes of
// the normal-mode coupling): host noise when given, else the two outputs of
// one Philox/Box-Muller block keyed by (m, a, i).
__device__ __forceinline__ void noise_pair(const MomentaArgs& m, int mp, int a, int i, double& g0, double& g1) {
    if (m.noise) {
        const size_t e = ((size_t)(2 * mp) * m.D + a) * m.N + i;
        g0 = m.noise[e];
        g1 = (2 * mp + 1 < m.P) ? m.noise[e + (size_t)m.D * m.N] : 0.0;
        return;
    }
    gauss2((uint32_t)(((size_t)mp * m.D + a) * m.N + i), (uint32_t)*m.step_ctr, m.ostep, m.replica, m.seed, g0, g1);
}

// Optional COM subtraction, optional half kick, optional O-step, optional
// partial COM sums of the result.  Grid (ceil(N/tc), D).
static __global__ void __launch_bounds__(256) k_momenta(MomentaArgs m, double* __restrict__ p) {
    extern __shared__ double smem[];
    const int P = m.P, N = m.N, D = m.D, TC = m.tc;
    double* pc = smem;                       // [P][TC]
    double* pn = smem + (size_t)P * TC;      // [P][TC]
    double* red = pn + (size_t)P * TC;       // [blockDim/WAVE]
    const int a = blockIdx.y;
    const int i0 = blockIdx.x * TC;
    const int tid = threadIdx.x;
    const int ncol = min(TC, N - i0);
    __shared__ double scm[4];
    if (m.com_in) block_com(m.com_in, m.ncom_in, D, N, P, scm);
    const double cm = m.com_in ? scm[a] : 0.0;
    for (int e = tid; e < P * TC; e += blockDim.x) {
        const int b = e / TC, c = e % TC;
        double v = 0.0;
        if (c < ncol) {
            const size_t g = ((size_t)b * D + a) * N + i0 + c;
            v = p[g];
            if (m.com_in) v -= cm;
            if (m.kick_f) v += m.half_dt * m.kick_f[g];
        }
        pc[e] = v;
    }
    if (m.stage == 1) {
        __syncthreads();
        for (int e = tid; e < P * TC; e += blockDim.x) {
            const int k = e / TC, c = e % TC;
            if (c < ncol) {
                double s = 0.0;
                for (int j = 0; j < P; ++j) s += m.Ct[(size_t)j * P + k] * pc[j * TC + c];
                m.modes[((size_t)k * D + a) * N + i0 + c] = s;
            }
        }
        return;
    }
    if (m.stage == 2) {
        __syncthreads();
        for (int e = tid; e < P * TC; e += blockDim.x) {
            const int k = e / TC, c = e % TC;
            pn[e] = c < ncol ? m.modes[((size_t)k * D + a) * N + i0 + c] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < P * TC; e += blockDim.x) {
            const int b = e / TC, c = e % TC;
            double s = 0.0;
            for (int k = 0; k < P; ++k) s += m.CTt[(size_t)k * P + b] * pn[k * TC + c];
            pc[e] = s;
        }
    } else if (m.thermostat == 1) {
        const int half = (P + 1) / 2;
        if (!m.nm) {
            for (int e = tid; e < half * TC; e += blockDim.x) {
                const int mp = e / TC, c = e % TC;
                if (c < ncol) {
                    double g[2];
                    noise_pair(m, mp, a, i0 + c, g[0], g[1]);
                    for (int h = 0; h < 2 && 2 * mp + h < P; ++h) {
                        double& v = pc[(2 * mp + h) * TC + c];
                        v = m.c1 * v + m.c2 * g[h];
                    }
                }
            }
        } else {
            __syncthreads();
            // noise of modes 2mp, 2mp+1 (one Box-Muller pair) into pn
            for (int e = tid; e < half * TC; e += blockDim.x) {
                const int mp = e / TC, c = e % TC;
                double g0 = 0.0, g1 = 0.0;
                if (c < ncol) noise_pair(m, mp, a, i0 + c, g0, g1);
                pn[2 * mp * TC + c] = g0;
                if (2 * mp + 1 < P) pn[(2 * mp + 1) * TC + c] = g1;
            }
            __syncthreads();
            // mode k of column c: sum_j C[k][j] p_j, then the Langevin update
            for (int e = tid; e < P * TC; e += blockDim.x) {
                const int k = e / TC, c = e % TC;
                double s = 0.0;
                for (int j = 0; j < P; ++j) s += m.Ct[(size_t)j * P + k] * pc[j * TC + c];
                pn[e] = (c < ncol) ? m.c1 * s + m.c2 * pn[e] : 0.0;
            }
            __syncthreads();
            for (int e = tid; e < P * TC; e += blockDim.x) {
                const int b = e / TC, c = e % TC;
                double s = 0.0;
                for (int k = 0; k < P; ++k) s += m.CTt[(size_t)k * P + b] * pn[k * TC + c];
                pc[e] = s;
            }
        }
    }
    __syncthreads();
    double part = 0.0;
    for (int e = tid; e < P * TC; e += blockDim.x) {
        const int b = e / TC, c = e % TC;
        if (c < ncol) {
            p[((size_t)b * D + a) * N + i0 + c] = pc[e];
            part += pc[e];
        }
    }
    if (m.com_out) {
        part = wave_sum(part);
        if (tid % WAVE == 0) red[tid / WAVE] = part;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < (int)blockDim.x / WAVE; ++w) s += red[w];
            m.com_out[(size_t)blockIdx.x * D + a] = s;
        }
    }
}

struct DriftArgs {
    int N, P, D;
    double half_dt, dt, mass;
    const double* com_in;
    int ncom_in;
    unsigned long long* step_ctr;   // incremented once per call when non-null
};

// p -= p_cm; p += dt/2 f; x += dt p / m  (velocity_verlet.cpp:24-38)
static __global__ void __launch_bounds__(256) k_kick_drift(DriftArgs d, double* __restrict__ x, double* __restrict__ p,
                                                    const double* __restrict__ f) {
    const size_t tot = (size_t)d.P * d.D * d.N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double scm[4];
    if (d.com_in) block_com(d.com_in, d.ncom_in, d.D, d.N, d.P, scm);
    if (e == 0 && d.step_ctr) *d.step_ctr += 1;
    if (e >= tot) return;
    const int a = (int)((e / d.N) % d.D);
    double pv = p[e];
    if (d.com_in) pv -= scm[a];
    pv += d.half_dt * f[e];
    p[e] = pv;
    x[e] += d.dt * pv / d.mass;
}

// p -= p_cm (the deferred zeroMomentum of the last O-step)
static __global__ void __launch_bounds__(256) k_com_subtract(int N, int P, int D, const double* __restrict__ com_in,
                                                      int ncom_in, double* __restrict__ p) {
    const size_t tot = (size_t)P * D * N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double scm[4];
    block_com(com_in, ncom_in, D, N, P, scm);
    if (e >= tot) return;
    const int a = (int)((e / N) % D);
    p[e] -= scm[a];
}

struct NmPropArgs {
    int N, P, D, tc;
    double half_dt, dt, mass;
    const double* Ct;       // C transposed, as in MomentaArgs
    const double* CTt;      // CT transposed
    const double* rot;      // [P][4]: cos(w dt), sin(w dt), m w, w
    const double* com_in;
    int ncom_in;
    const double* fext;     // external forces of the previous evaluation
    unsigned long long* step_ctr;
};

// Normal-mode propagator for distinguishable particles: external half kick,
// exact free ring-polymer evolution in normal modes, back to Cartesian
// (normal_modes_propagator.cpp:19-53).  Grid (ceil(N/tc), D).
static __global__ void __launch_bounds__(256) k_nm_propagate(NmPropArgs m, double* __restrict__ x, double* __restrict__ p) {
    extern __shared__ double smem[];
    const int P = m.P, N = m.N, D = m.D, TC = m.tc;
    double* xc = smem;
    double* pc = xc + (size_t)P * TC;
    double* qn = pc + (size_t)P * TC;
    double* pn = qn + (size_t)P * TC;
    const int a = blockIdx.y;
    const int i0 = blockIdx.x * TC;
    const int tid = threadIdx.x;
    const int ncol = min(TC, N - i0);
    if (tid == 0 && blockIdx.x == 0 && blockIdx.y == 0 && m.step_ctr) *m.step_ctr += 1;
    __shared__ double scm[4];
    if (m.com_in) block_com(m.com_in, m.ncom_in, D, N, P, scm);
    const double cm = m.com_in ? scm[a] : 0.0;
    for (int e = tid; e < P * TC; e += blockDim.x) {
        const int b = e / TC, c = e % TC;
        double xv = 0.0, pv = 0.0;
        if (c < ncol) {
            const size_t g = ((size_t)b * D + a) * N + i0 + c;
            xv = x[g];
            pv = p[g];
            if (m.com_in) pv -= cm;
            pv += m.half_dt * m.fext[g];
        }
        xc[e] = xv;
        pc[e] = pv;
    }
    __syncthreads();
    for (int e = tid; e < P * TC; e += blockDim.x) {
        const int k = e / TC, c = e % TC;
        double q = 0.0, pp = 0.0;
        for (int j = 0; j < P; ++j) {
            const double cj = m.Ct[(size_t)j * P + k];
            q += cj * xc[j * TC + c];
            pp += cj * pc[j * TC + c];
        }
        const double cs = m.rot[4 * k], sn = m.rot[4 * k + 1], mw = m.rot[4 * k + 2], w = m.rot[4 * k + 3];
        if (w == 0) {
            qn[e] = q + m.dt / m.mass * pp;
            pn[e] = pp;
        } else {
            qn[e] = cs * q + sn / mw * pp;
            pn[e] = (-1) * mw * sn * q + cs * pp;
        }
    }
    __syncthreads();
    for (int e = tid; e < P * TC; e += blockDim.x) {
        const int b = e / TC, c = e % TC;
        if (c >= ncol) continue;
        double xv = 0.0, pv = 0.0;
        for (int k = 0; k < P; ++k) {
            const double ck = m.CTt[(size_t)k * P + b];
            xv += ck * qn[k * TC + c];
            pv += ck * pn[k * TC + c];
        }
        const size_t g = ((size_t)b * D + a) * N + i0 + c;
        x[g] = xv;
        p[g] = pv;
    }
}

}  // namespace pimdb
