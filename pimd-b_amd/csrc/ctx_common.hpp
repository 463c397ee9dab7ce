// ctx_common.hpp -- host-side helpers shared by the single-system context,
// the exchange plan and the batched replicas: derived constants, parameter
// validation, normal-mode tables, layout transposes and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "pimdb_gpu.h"
#include "exchange.hpp"
#include "forces.hpp"
#include "observables.hpp"

namespace pimdb {

// Message of the last failed create call on this thread.
inline std::string& create_error_slot() {
    static thread_local std::string msg;
    return msg;
}
inline void set_create_error(const std::string& m) { create_error_slot() = m; }
inline const char* create_error() { return create_error_slot().c_str(); }

// Constants derived as in Simulation::Simulation (src/simulation.cpp:37-52),
// BosonicExchangeBase (bosonic_exchange_base.cpp:16-18), LangevinThermostat
// (langevin.cpp:10-13), the potential factory (:612-645) and the cutoff rule
// (:83-90).  kB = hbar = 1 (include/units.h:18-22).
struct Derived {
    double beta, thermo_beta, omega_p, k, bt, c1, c2, cutoff, int_k;
    ExtPot ext;
    bool bosonic;
};

inline Derived derive(const pimdb_params& p) {
    Derived d{};
    d.beta = 1.0 / (1.0 * p.temperature);
    d.thermo_beta = d.beta / p.nbeads;
    d.omega_p = p.nbeads / (d.beta * 1.0);
    d.k = p.mass * d.omega_p * d.omega_p;
    d.bt = d.beta / p.nbeads;
    d.c1 = std::exp(-0.5 * p.gamma * p.dt);
    d.c2 = std::sqrt((1 - d.c1 * d.c1) * p.mass / d.thermo_beta);
    double c = (p.int_kind == PIMDB_POT_FREE) ? 0.0 : p.cutoff;
    if (p.pbc) c = std::min(c, 0.5 * p.size);
    d.cutoff = c;
    d.ext = ExtPot{p.ext_kind, 0.0, 0.0, 0.0, 0.0};
    if (p.ext_kind == PIMDB_POT_HARMONIC) {
        d.ext.a = p.mass * p.ext_omega * p.ext_omega;
    } else if (p.ext_kind == PIMDB_POT_DOUBLE_WELL) {
        d.ext.a = 4 * p.mass * p.ext_strength;
        d.ext.b = p.ext_location * p.ext_location;
        d.ext.c = p.mass * p.ext_strength;
    } else if (p.ext_kind == PIMDB_POT_COSINE) {
        const double kc = 2 * 3.14159265358979323846 / p.size;
        d.ext.a = -p.ext_amplitude * kc;
        d.ext.b = kc;
        d.ext.c = p.ext_phase;
        d.ext.d = p.ext_amplitude;
    }
    d.int_k = p.mass * p.int_omega * p.int_omega;
    d.bosonic = p.bosonic && p.nbeads > 1;   // simulation.cpp:690
    return d;
}

// Checks of src/params.cpp:14-247 that concern the hot path.
inline bool validate_params(const pimdb_params& p, std::string* why) {
    auto bad = [&](const char* m) { if (why) *why = m; return false; };
    if (p.natoms < 1) return bad("The specified number of particles is smaller than one");
    if (p.nbeads < 1) return bad("The specified number of beads is less than one");
    if (p.ndim < 1 || p.ndim > 3) return bad("ndim must be 1, 2 or 3");
    if (!(p.mass > 0.0)) return bad("The provided mass is unphysical");
    if (!(p.temperature > 0.0)) return bad("The specified temperature is unphysical");
    if (!(p.size > 0.0)) return bad("The provided system size is unphysical");
    if (p.int_kind < 0 || p.int_kind > 3) return bad("The specified interaction potential is not supported!");
    if (p.ext_kind != PIMDB_POT_FREE && p.ext_kind != PIMDB_POT_HARMONIC && p.ext_kind != PIMDB_POT_DOUBLE_WELL &&
        p.ext_kind != PIMDB_POT_COSINE)
        return bad("The specified external potential is not supported!");
    if (p.propagator != PIMDB_PROP_CARTESIAN && p.propagator != PIMDB_PROP_NORMAL_MODES) return bad("unknown propagator");
    if (p.bosonic && p.propagator == PIMDB_PROP_NORMAL_MODES)
        return bad("Normal modes propogation is currently not available for bosons!");
    if (p.thermostat < PIMDB_THERMO_NONE || p.thermostat > PIMDB_THERMO_NOSE_HOOVER_NP_DIM) return bad("unknown thermostat");
    // nose_hoover.cpp:141-143 reads eta_dot[nchains - 2] of the chain, outside
    // it for nchains = 1, so chains of one component are not supported
    if (p.thermostat >= PIMDB_THERMO_NOSE_HOOVER && p.nchains < 2)
        return bad("Nose-Hoover chains need nchains >= 2");
    if (p.nmthermostat && p.thermostat == PIMDB_THERMO_NONE) return bad("nmthermostat cannot be used in nve ensemble!");
    if (p.nbeads > 2048) return bad("nbeads above 2048 is not supported");
    if (p.exchange_algo != PIMDB_EXCHANGE_QUADRATIC && p.exchange_algo != PIMDB_EXCHANGE_FACTORIAL)
        return bad("unknown bosonic exchange algorithm");
    if (p.exchange_algo == PIMDB_EXCHANGE_FACTORIAL && p.bosonic && p.natoms > 10)
        return bad("the factorial exchange enumerates natoms! permutations: natoms <= 10");
    return true;
}

// Particles per tile of the column kernels (all P beads of tc particles):
// about one output per thread of a 256-thread workgroup.
inline int column_tile(int P) { return std::max(1, std::min(64, 256 / std::max(1, P))); }

// Transposes of the normal-mode matrices for coalesced reads:
// CTr[j][k] = C[k][j], CTT[k][b] = CT[b][k].
inline void transpose_square(int P, const double* A, double* At) {
    for (int r = 0; r < P; ++r)
        for (int c = 0; c < P; ++c) At[(size_t)c * P + r] = A[(size_t)r * P + c];
}

// Cartesian <-> normal-mode matrices exactly as NormalModes builds them
// (src/normal_modes.cpp:47-81): C[k][j] is row k as rank k computes it,
// CT[b][k] is the nm->cart row of rank b.
inline void nm_matrices(int P, double* C, double* CT) {
    const double pi = 3.14159265358979323846;
    for (int b = 0; b < P; ++b) {
        double pref;
        const double fund_freq = 2 * pi / P * b;
        double* row = C + (size_t)b * P;
        if (b == 0) {
            pref = 1 / std::sqrt((double)P);
            for (int i = 0; i < P; ++i) row[i] = pref;
        } else if (b < 0.5 * P) {
            pref = std::sqrt(2.0 / P);
            for (int i = 0; i < P; ++i) row[i] = pref * std::cos(fund_freq * i);
        } else if (b == 0.5 * P) {
            pref = 1 / std::sqrt((double)P);
            for (int i = 0; i < P; ++i) row[i] = pref * (i % 2 == 0 ? 1.0 : -1.0);
        } else {
            pref = std::sqrt(2.0 / P);
            for (int i = 0; i < P; ++i) row[i] = -pref * std::sin(fund_freq * i);
        }
        double* crow = CT + (size_t)b * P;
        pref = std::sqrt(2.0 / P);
        for (int i = 0; i < P; ++i) crow[i] = 0.0;
        crow[0] = 1 / std::sqrt((double)P);
        for (int i = 1; i < 0.5 * P; ++i) crow[i] = pref * std::cos(fund_freq * i);
        if (P % 2 == 0) crow[P / 2] = 1 / std::sqrt((double)P) * (b % 2 == 0 ? 1.0 : -1.0);
        for (int i = (int)std::ceil(0.5 * (P + 1)); i < P; ++i) crow[i] = -pref * std::sin(fund_freq * i);
    }
}

// Free ring-polymer rotation per mode (normal_modes_propagator.cpp:12-16):
// rot[4k..4k+3] = cos(w dt), sin(w dt), m w, w with w = 2 w_P sin(k pi / P).
inline void nm_rotation(const pimdb_params& p, const Derived& d, double* rot) {
    const double pi = 3.14159265358979323846;
    for (int b = 0; b < p.nbeads; ++b) {
        const double freq = 2 * d.omega_p * std::sin(b * pi / p.nbeads);
        rot[4 * b] = std::cos(freq * p.dt);
        rot[4 * b + 1] = std::sin(freq * p.dt);
        rot[4 * b + 2] = p.mass * freq;
        rot[4 * b + 3] = freq;
    }
}

// [b][i][a] (host, stack of dVec) <-> [b][a][i] (device SoA)
inline void to_soa(const double* h, double* s, int N, int P, int D) {
    for (int b = 0; b < P; ++b)
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) s[((size_t)b * D + a) * N + i] = h[((size_t)b * N + i) * D + a];
}
inline void from_soa(const double* s, double* h, int N, int P, int D) {
    for (int b = 0; b < P; ++b)
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) h[((size_t)b * N + i) * D + a] = s[((size_t)b * D + a) * N + i];
}

// The smallest double t with sqrt(t) >= c (IEEE sqrt, correctly rounded), so
// that r2 < t exactly when sqrt(r2) < c: the reference's strict cutoff test
// on the norm (simulation.cpp:444) as one comparison of the squared distance.
inline double cutoff_threshold(double c) {
    if (!(c > 0.0)) return 0.0;
    double t = c * c;
    while (std::sqrt(t) < c) t = std::nextafter(t, INFINITY);
    for (;;) {
        const double lower = std::nextafter(t, 0.0);
        if (std::sqrt(lower) >= c) t = lower;
        else break;
    }
    return t;
}

inline ForceArgs make_force_args(const pimdb_params& p, const Derived& d) {
    ForceArgs fa{};
    fa.N = p.natoms; fa.P = p.nbeads; fa.bosonic = d.bosonic ? 1 : 0; fa.pbc = p.pbc;
    fa.L = p.size; fa.k_spring = d.k;
    fa.ext = d.ext;
    fa.int_k = d.int_k; fa.strength = p.int_strength;
    fa.cutoff = d.cutoff;
    fa.cut2 = cutoff_threshold(d.cutoff);
    return fa;
}

// LDS of k_exchange_recursions (1024 threads per direction).
inline size_t rec_lds_bytes(int N, int D) { return sizeof(double) * rec_shared_doubles(N, D, 1024 / WAVE); }

inline size_t device_lds_limit() {
    static size_t lim = 0;
    if (!lim) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            lim = prop.sharedMemPerBlock;
        if (!lim) lim = 65536;
    }
    return lim;
}

inline void launch_cycle_energies(int D, int N, const double* x1, const double* xP, double k, int pbc, double L,
                                  double* E, double* ET, hipStream_t s) {
    dim3 grid((N + 3) / 4), block(256);
    switch (D) {
    case 1: hipLaunchKernelGGL(k_cycle_energies<1>, grid, block, 0, s, N, x1, xP, k, pbc, L, E, ET); break;
    case 2: hipLaunchKernelGGL(k_cycle_energies<2>, grid, block, 0, s, N, x1, xP, k, pbc, L, E, ET); break;
    default: hipLaunchKernelGGL(k_cycle_energies<3>, grid, block, 0, s, N, x1, xP, k, pbc, L, E, ET); break;
    }
}

inline void launch_bead_sums(int D, int N, int P, double k, int pbc, double L, const double* x, const double* p,
                             double* out, hipStream_t s) {
    switch (D) {
    case 1: hipLaunchKernelGGL(k_bead_sums<1>, dim3(P), dim3(256), 0, s, N, P, k, pbc, L, x, p, out); break;
    case 2: hipLaunchKernelGGL(k_bead_sums<2>, dim3(P), dim3(256), 0, s, N, P, k, pbc, L, x, p, out); break;
    default: hipLaunchKernelGGL(k_bead_sums<3>, dim3(P), dim3(256), 0, s, N, P, k, pbc, L, x, p, out); break;
    }
}

template <int D>
inline void launch_bead_energies_d(const pimdb_params& p, const Derived& d, const double* x, double* part,
                                   hipStream_t s) {
    const ForceArgs fa = make_force_args(p, d);
    dim3 grid((p.natoms + FORCE_TILE - 1) / FORCE_TILE, p.nbeads);
    const size_t red = sizeof(double) * 3 * FORCE_WAVES;
    const size_t lds = sizeof(double) * (size_t)D * p.natoms + red;
    const dim3 blk(FORCE_TILE * FORCE_WAVES);
    if (lds > device_lds_limit()) {   // the slice read in place
        switch (p.int_kind) {
        case POT_HARMONIC: hipLaunchKernelGGL((k_bead_energies<D, POT_HARMONIC, true>), grid, blk, red, s, fa, x, part); break;
        case POT_DIPOLE: hipLaunchKernelGGL((k_bead_energies<D, POT_DIPOLE, true>), grid, blk, red, s, fa, x, part); break;
        case POT_AZIZ: hipLaunchKernelGGL((k_bead_energies<D, POT_AZIZ, true>), grid, blk, red, s, fa, x, part); break;
        default: hipLaunchKernelGGL((k_bead_energies<D, POT_FREE, true>), grid, blk, red, s, fa, x, part); break;
        }
        return;
    }
    switch (p.int_kind) {
    case POT_HARMONIC: hipLaunchKernelGGL((k_bead_energies<D, POT_HARMONIC>), grid, blk, lds, s, fa, x, part); break;
    case POT_DIPOLE: hipLaunchKernelGGL((k_bead_energies<D, POT_DIPOLE>), grid, blk, lds, s, fa, x, part); break;
    case POT_AZIZ: hipLaunchKernelGGL((k_bead_energies<D, POT_AZIZ>), grid, blk, lds, s, fa, x, part); break;
    default: hipLaunchKernelGGL((k_bead_energies<D, POT_FREE>), grid, blk, lds, s, fa, x, part); break;
    }
}

inline void launch_bead_energies(int D, const pimdb_params& p, const Derived& d, const double* x, double* part,
                                 hipStream_t s) {
    switch (D) {
    case 1: launch_bead_energies_d<1>(p, d, x, part, s); break;
    case 2: launch_bead_energies_d<2>(p, d, x, part, s); break;
    default: launch_bead_energies_d<3>(p, d, x, part, s); break;
    }
}

// Per-bead sums -> the observables the reference's ObservablesLogger prints
// after its MPI_Allreduce over beads (observable.cpp:92-116), in atomic units.
inline void combine_observables(const pimdb_params& p, const Derived& d, const double* sums, const double* epart,
                                int nt, const double* prim, double VN, double* out) {
    const int N = p.natoms, P = p.nbeads, D = p.ndim;
    const bool ext_free = p.ext_kind == PIMDB_POT_FREE;
    const bool int_free = p.int_kind == PIMDB_POT_FREE;
    double kinetic = 0, potential = 0, ext_pot = 0, int_pot = 0, virial = 0, temperature = 0, cl_kin = 0,
           cl_spring = 0;
    for (int b = 0; b < P; ++b) {
        const double spring_b = sums[2 * b], p2 = sums[2 * b + 1];
        double kin = 0.5 * D * N / d.beta;
        if (b == 0 && d.bosonic) kin += prim[0];
        else kin -= spring_b / P;
        kinetic += kin;
        double e = 0, ip = 0, v = 0;
        for (int t = 0; t < nt; ++t) {
            const double* o = epart + ((size_t)b * nt + t) * 3;
            e += o[0]; ip += o[1]; v += o[2];
        }
        if (!ext_free && !int_free) {
            ext_pot += e / P;
            int_pot += ip / P;
        }
        if (!ext_free || !int_free) {
            potential += (e + ip) / P;
            virial += v * (0.5 / P);
        }
        const double ke = p2 * (0.5 / p.mass);
        cl_kin += ke;
        temperature += 2.0 * ke / ((double)D * N * P) / P;
        cl_spring += (b == 0 && d.bosonic) ? VN : spring_b;
    }
    out[0] = kinetic; out[1] = potential; out[2] = ext_pot; out[3] = int_pot; out[4] = virial;
    out[5] = temperature; out[6] = cl_kin; out[7] = cl_spring;
    out[8] = d.bosonic ? prim[1] : 0.0;
    out[9] = d.bosonic ? prim[2] : 0.0;
}

}  // namespace pimdb
