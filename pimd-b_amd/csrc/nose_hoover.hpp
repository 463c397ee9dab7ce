// nose_hoover.hpp -- Nose-Hoover chain thermostats on the GPU (the plain
// form): one chain per bead, per particle of each bead, or per degree of
// freedom of each bead, as the reference's one thermostat object per bead
// (rank) holds them.
//
// Reference: src/thermostats/nose_hoover.cpp (Tuckerman et al. 2006 and its
// LAMMPS implementation).  For a chain of nc components (eta, eta_dot,
// eta_dot_dot) and the current kinetic measure E = sum p^2 / m of its degrees
// of freedom, singleChainStep (:96-150) updates the chain by a quarter step,
// returns the momentum scale exp(-dt/2 eta_dot[0]) and updates the chain by
// another quarter step with the rescaled E.  Masses Q_i = beta / P (hbar = 1),
// Q_1 = ndof Q_i; target ndof / beta~ (:10-32, 152-163, 198-208).
//
// The O-step for Cartesian coupling is one kernel with one workgroup per
// bead: the pending zeroMomentum, the optional half kick, the chain steps and
// the rescaling, and the partial momentum sums for the next zeroMomentum.  For
// normal-mode coupling the same kernel runs on the modes (written and read
// back by the normal-mode transform kernels, k_momenta stages 1 and 2).
#pragma once

#include "device_common.hpp"
#include "integrator.hpp"
#include "pimdb_gpu.h"

namespace pimdb {

struct NhArgs {
    int N, P, D, nc;
    int kind;                    // PIMDB_THERMO_NOSE_HOOVER, _NP or _NP_DIM
    int groups;                  // chains per bead: 1, N or N*D
    double mass, dt, Q1, Qi, required, thermo_beta;
    const double* com_in;        // partial momentum sums to subtract first, or null
    int ncom_in;
    const double* kick_f;        // forces for a half kick first, or null
    double half_dt;
    double* com_out;             // [P][D] partial sums of the result, or null
    double *eta, *eta_dot, *eta_dot_dot;   // [P][groups][nc]
};

// singleChainStep on one chain; returns the momentum scale.
__device__ inline double nh_chain_step(const NhArgs& h, double* e, double* ed, double* edd, double current_energy) {
    const int nc = h.nc;
    const double dt2 = 0.5 * h.dt, dt4 = 0.25 * h.dt, dt8 = 0.125 * h.dt;
    double exp_factor = 0.0;
    edd[0] = (current_energy - h.required) / h.Q1;
    ed[nc - 1] += edd[nc - 1] * dt4;
    for (int i = nc - 2; i >= 0; i--) {
        exp_factor = exp(-dt8 * ed[i + 1]);
        ed[i] *= exp_factor;
        ed[i] += edd[i] * dt4;
        ed[i] *= exp_factor;
    }
    const double scale = exp(-dt2 * ed[0]);
    for (int i = 0; i < nc; i++) e[i] += dt2 * ed[i];
    edd[0] = (current_energy * scale * scale - h.required) / h.Q1;
    ed[0] *= exp_factor;
    ed[0] += edd[0] * dt4;
    ed[0] *= exp_factor;
    double Q_former = h.Q1;
    for (int i = 1; i < nc - 1; i++) {
        exp_factor = exp(-dt8 * ed[i + 1]);
        ed[i] *= exp_factor;
        edd[i] = (Q_former * ed[i - 1] * ed[i - 1] - 1 / h.thermo_beta) / h.Qi;
        ed[i] += edd[i] * dt4;
        ed[i] *= exp_factor;
        Q_former = h.Qi;
    }
    edd[nc - 1] = (h.Qi * ed[nc - 2] * ed[nc - 2] - 1 / h.thermo_beta) / h.Qi;
    ed[nc - 1] += edd[nc - 1] * dt4;
    return scale;
}

// One Nose-Hoover O-step on p [b][a][i] (Cartesian momenta, or the modes
// under normal-mode coupling).  Grid: P workgroups (bead b = blockIdx.x).
static __global__ void __launch_bounds__(256) k_nh_ostep(NhArgs h, double* __restrict__ p) {
    const int b = blockIdx.x, N = h.N, D = h.D, nc = h.nc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    __shared__ double scm[4];
    __shared__ double red[256 / WAVE];
    __shared__ double sscale;
    if (h.com_in) block_com(h.com_in, h.ncom_in, D, N, h.P, scm);
    double* pb = p + (size_t)b * D * N;
    const size_t cbase = (size_t)b * h.groups * nc;
    // the pending zeroMomentum and the half kick, in place
    if (h.com_in || h.kick_f) {
        for (int e = tid; e < D * N; e += nthr) {
            double v = pb[e];
            if (h.com_in) v -= scm[e / N];
            if (h.kick_f) v += h.half_dt * h.kick_f[(size_t)b * D * N + e];
            pb[e] = v;
        }
        __syncthreads();
    }
    if (h.kind == PIMDB_THERMO_NOSE_HOOVER) {
        double s = 0.0;
        for (int e = tid; e < D * N; e += nthr) s += pb[e] * pb[e];
        s = wave_sum(s);
        if (tid % WAVE == 0) red[tid / WAVE] = s;
        __syncthreads();
        if (tid == 0) {
            double ce = 0.0;
            for (int w = 0; w < nthr / WAVE; ++w) ce += red[w];
            ce /= h.mass;
            sscale = nh_chain_step(h, h.eta + cbase, h.eta_dot + cbase, h.eta_dot_dot + cbase, ce);
        }
        __syncthreads();
        const double sc = sscale;
        for (int e = tid; e < D * N; e += nthr) pb[e] = pb[e] * sc;
    } else if (h.kind == PIMDB_THERMO_NOSE_HOOVER_NP) {
        for (int i = tid; i < N; i += nthr) {
            double ce = 0.0;
            for (int a = 0; a < D; ++a) ce += pb[a * N + i] * pb[a * N + i];
            ce /= h.mass;
            const size_t o = cbase + (size_t)i * nc;
            const double sc = nh_chain_step(h, h.eta + o, h.eta_dot + o, h.eta_dot_dot + o, ce);
            for (int a = 0; a < D; ++a) pb[a * N + i] = pb[a * N + i] * sc;
        }
    } else {
        for (int e = tid; e < D * N; e += nthr) {
            const int a = e / N, i = e % N;
            const double ce = pb[e] * pb[e] / h.mass;
            const size_t o = cbase + ((size_t)i * D + a) * nc;
            const double sc = nh_chain_step(h, h.eta + o, h.eta_dot + o, h.eta_dot_dot + o, ce);
            pb[e] = pb[e] * sc;
        }
    }
    if (h.com_out) {
        __syncthreads();
        if (tid < WAVE * D) {
            const int a = tid / WAVE, lane = tid % WAVE;
            double s = 0.0;
            for (int i = lane; i < N; i += WAVE) s += pb[a * N + i];
            s = wave_sum(s);
            if (lane == 0) h.com_out[(size_t)b * D + a] = s;
        }
    }
}

// getAdditionToH summed over all chains of all beads (nose_hoover.cpp:41-66,
// 186-192, 229-237; the nh_energy column).  One workgroup.
static __global__ void __launch_bounds__(256) k_nh_energy(NhArgs h, int ndof, double* __restrict__ out) {
    __shared__ double red[256 / WAVE];
    const int nchain = h.P * h.groups, nc = h.nc;
    double s = 0.0;
    for (int c = threadIdx.x; c < nchain; c += blockDim.x) {
        const double* e = h.eta + (size_t)c * nc;
        const double* ed = h.eta_dot + (size_t)c * nc;
        double a = 0.5 * h.Q1 * ed[0] * ed[0];
        a += ndof * e[0] / h.thermo_beta;
        for (int i = 1; i < nc; i++) {
            a += 0.5 * h.Qi * ed[i] * ed[i];
            a += e[i] / h.thermo_beta;
        }
        s += a;
    }
    s = wave_sum(s);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)blockDim.x / WAVE; ++w) t += red[w];
        *out = t;
    }
}

}  // namespace pimdb
