// forces.hpp -- per-slice physical forces (external trap + pair interaction
// with cutoff and minimum image) and the interior ring-polymer springs.
//
// Reference: Simulation::updateForces / updateSpringForces /
// updatePhysicalForces, src/simulation.cpp:353-455; potentials in
// src/potentials/{harmonic,dipole,aziz}.cpp.
//
// One workgroup handles 64 particles i of one bead (time slice) b: the bead's
// N x D positions are staged in LDS, each of the 16 waves runs a sixteenth of
// the partners j (four waves per SIMD hide the FP64 latency chains), and the
// partial forces are reduced in LDS.  Every ordered pair
// (i,j) evaluates gradV(r_i - r_j); for central potentials this is bitwise
// the negation of gradV(r_j - r_i), so the result equals the reference's
// Newton's-third-law loop up to summation order.
#pragma once

#include "device_common.hpp"
#include "pairs.hpp"

namespace pimdb {

struct ForceArgs {
    int N, P, bosonic, pbc;
    double L, k_spring;
    ExtPot ext;              // external potential
    double int_k, strength;  // pair harmonic constant, dipole strength
    double cutoff;           // effective cutoff (0: no pair loop; < 0: all pairs)
    double cut2;             // cutoff_threshold(cutoff)
};

constexpr int FORCE_TILE = 64;
constexpr int FORCE_WAVES = 16;

// GX: the slice is read in place from device memory instead of being staged
// in LDS (slices larger than a workgroup's LDS); the LDS then holds only the
// reduction.
template <int D, int KIND, bool GX = false>
static __global__ void __launch_bounds__(FORCE_TILE * FORCE_WAVES)
k_bead_forces(ForceArgs fa, const double* __restrict__ x, double* __restrict__ f, double* __restrict__ fext) {
    extern __shared__ double smem[];
    double* red = smem + (GX ? 0 : (size_t)D * fa.N);   // [FORCE_WAVES-1][D][FORCE_TILE]
    const int N = fa.N, P = fa.P;
    const int b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const size_t bs = (size_t)D * N;
    const double* xb = x + bs * b;
    const double* sx;                                   // [D][N]
    if constexpr (GX) {
        sx = xb;
    } else {
        for (int e = tid; e < D * N; e += blockDim.x) smem[e] = xb[e];
        __syncthreads();
        sx = smem;
    }

    const int i = blockIdx.x * FORCE_TILE + lane;
    const bool pbc = fa.pbc != 0;
    double acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0;
    if constexpr (KIND != POT_FREE) {
        if (i < N && fa.cutoff != 0.0) {
            double xi[D];
#pragma unroll
            for (int a = 0; a < D; ++a) xi[a] = sx[a * N + i];
            const int chunk = (N + FORCE_WAVES - 1) / FORCE_WAVES;
            const int jlo = wave * chunk, jhi = min(N, jlo + chunk);
            const PairParams pp = make_pair_params(fa.L, fa.pbc, fa.cutoff, fa.cut2, fa.int_k, fa.strength);
            auto run = [&](auto img) {
                constexpr int IMG = decltype(img)::value;
                for (int j = jlo; j < jhi; ++j) {
                    if (j == i) continue;
                    double xj[D], d[D], g;
#pragma unroll
                    for (int a = 0; a < D; ++a) xj[a] = sx[a * N + j];
                    if (pair_grad<D, KIND, IMG>(xi, xj, pp, d, g)) {
#pragma unroll
                        for (int a = 0; a < D; ++a) acc[a] = fma(-g, d[a], acc[a]);
                    }
                }
            };
            switch (pair_image(pp)) {
            case 0: run(std::integral_constant<int, 0>()); break;
            case 1: run(std::integral_constant<int, 1>()); break;
            default: run(std::integral_constant<int, 2>()); break;
            }
        }
        if (wave > 0) {
#pragma unroll
            for (int a = 0; a < D; ++a) red[((wave - 1) * D + a) * FORCE_TILE + lane] = acc[a];
        }
        __syncthreads();
    }
    if (wave != 0 || i >= N) return;
    if constexpr (KIND != POT_FREE) {
        for (int w = 1; w < FORCE_WAVES; ++w) {
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] += red[((w - 1) * D + a) * FORCE_TILE + lane];
        }
    }
    const bool exchange_bead = fa.bosonic && (b == 0 || b == P - 1);
    const double* xprev = x + bs * ((b - 1 + P) % P);
    const double* xnext = x + bs * ((b + 1) % P);
    double xv[D], fe[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xv[a] = sx[a * N + i];
    ext_force<D>(fa.ext, xv, fe);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const size_t idx = bs * b + (size_t)a * N + i;
        const double xi = xv[a];
        const double phys = fe[a] + acc[a];
        if (fext) fext[idx] = phys;
        double spring = 0.0;
        if (!exchange_bead) {
            const double dp = disp(xi, xprev[(size_t)a * N + i], pbc, fa.L);
            const double dn = disp(xi, xnext[(size_t)a * N + i], pbc, fa.L);
            spring = fa.k_spring * (dp + dn);
        }
        // exchange beads get their spring part added by k_exchange_forces
        f[idx] = exchange_bead ? phys : spring + phys;
    }
}

// Pair and external potential energies and virial per bead, for the energy
// observable (src/observables/energy.cpp:56-111): out[b] = {ext_pot, int_pot,
// virial_sum} (un-normalised sums of the reference).
// GX as for k_bead_forces.
template <int D, int KIND, bool GX = false>
static __global__ void __launch_bounds__(FORCE_TILE * FORCE_WAVES)
k_bead_energies(ForceArgs fa, const double* __restrict__ x, double* __restrict__ part) {
    extern __shared__ double smem[];
    const int N = fa.N;
    const int b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const size_t bs = (size_t)D * N;
    const double* xb = x + bs * b;
    const double* sx;
    if constexpr (GX) {
        sx = xb;
    } else {
        for (int e = tid; e < D * N; e += blockDim.x) smem[e] = xb[e];
        __syncthreads();
        sx = smem;
    }
    const int i = blockIdx.x * FORCE_TILE + lane;
    const bool pbc = fa.pbc != 0;
    double ext = 0.0, ipot = 0.0, vir = 0.0;
    if (i < N) {
        double xi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) xi[a] = sx[a * N + i];
        if (wave == 0 && fa.ext.kind != POT_FREE) {
            double fe[D];
            ext = ext_energy<D>(fa.ext, xi);
            ext_force<D>(fa.ext, xi, fe);
#pragma unroll
            for (int a = 0; a < D; ++a) vir -= xi[a] * fe[a];
        }
        if constexpr (KIND != POT_FREE) {
            if (fa.cutoff != 0.0) {
                // pairs (i, j > i): energy.cpp:76-94, virial uses r_i only (i < j)
                const int chunk = (N + FORCE_WAVES - 1) / FORCE_WAVES;
                const int jlo = max(i + 1, wave * chunk), jhi = min(N, wave * chunk + chunk);
                for (int j = jlo; j < jhi; ++j) {
                    double d[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) d[a] = disp(sx[a * N + j], xi[a], pbc, fa.L);
                    const double r2 = norm2<D>(d);
                    const double norm = sqrt(r2);
                    if (norm < fa.cutoff || fa.cutoff < 0.0) {
                        const double g = pair_grad_factor<KIND>(norm, fa.int_k, fa.strength);
                        ipot += pair_energy<KIND>(norm, r2, fa.int_k, fa.strength);
#pragma unroll
                        for (int a = 0; a < D; ++a) vir -= xi[a] * (-(g * d[a]));
                    }
                }
            }
        }
    }
    ext = wave_sum(ext);
    ipot = wave_sum(ipot);
    vir = wave_sum(vir);
    double* red = smem + (GX ? 0 : (size_t)D * N);
    if (lane == 0) {
        red[wave * 3 + 0] = ext;
        red[wave * 3 + 1] = ipot;
        red[wave * 3 + 2] = vir;
    }
    __syncthreads();
    if (tid == 0) {
        double e = 0, ip = 0, v = 0;
        for (int w = 0; w < FORCE_WAVES; ++w) { e += red[w * 3]; ip += red[w * 3 + 1]; v += red[w * 3 + 2]; }
        double* o = part + ((size_t)b * gridDim.x + blockIdx.x) * 3;
        o[0] = e; o[1] = ip; o[2] = v;
    }
}

}  // namespace pimdb
