// batch.hpp -- the batched persistent form: many independent ring polymers
// (uneven N, P, T, potentials' parameters) advanced by ONE kernel launch per
// MD step.  Each workgroup owns whole replicas, taken from a static schedule
// (longest-processing-time first, computed on the host), and runs the full
// OBABO step of a replica with workgroup barriers only:
//
//   O-step (+ zeroMomentum) -> B + A -> pair/ext/spring forces of all beads
//   -> forward recursion -> backward recursion with the exchange forces on
//   beads 1 and P -> B -> O-step (+ zeroMomentum)
//
// (Simulation::run, src/simulation.cpp:246-259).  The device noise uses the
// same Philox keys as the single-system form, so a one-replica batch follows
// the single-system trajectory up to summation order.
#pragma once

#include "device_common.hpp"
#include "exchange_batch.hpp"
#include "nose_hoover.hpp"
#include "pairs.hpp"

namespace pimdb {

struct RepDesc {
    int N, P, bosonic, pbc;
    int thermostat, nm, fixcom;
    int plds;                   // pair forces accumulated in LDS (LDS permitting)
    int fused;                  // normal-mode O-step on the matrix cores with B + A fused in (see wg_momenta)
    int xslots;                 // exchange chunk slots that the workgroup's LDS holds (1 .. waves)
    double L, invL, mass, dt;
    double k, bt, c1, c2;
    double int_k, strength, cutoff, cut2;
    ExtPot ext;
    uint32_t seed, rid;
    unsigned long long off_x;   // P*D*N block of x, p, f
    unsigned long long off_V;   // N+1 block of V
    unsigned long long off_C;   // P*P block of Ct, CTt
    // The fields below are read only by the normal-mode-propagator and Nose-Hoover paths (the EXT kernels);
    // their arrays are reached from here rather than from BatchPtrs, whose copy in every other kernel's
    // frame stays as it is.
    int nmprop;                 // normal-mode propagator (distinguishable particles): the kicks take the
                                // physical forces fx, the drift is wg_nm_drift's
    union {
        double* fx;             // nmprop: this replica's physical forces [P][D][N]
        // Bosonic replicas (never nmprop: the normal-mode propagator is for distinguishable particles): the
        // exchange's workspace in device memory (xb_layout with one chunk slot per wave) when the layout does
        // not fit the workgroup's LDS share, else null; a batch with such a replica runs the EXT kernels.
        double* xws;
    };
    const double* rot;          // nmprop: [P][4] cos(w dt), sin(w dt), m w, w of each mode
    // Nose-Hoover chains (thermostat >= PIMDB_THERMO_NOSE_HOOVER, Cartesian coupling): chain length, masses
    // Q_1 and Q_i, target ndof / beta~, beta~ (nose_hoover.cpp:10-32), and the replica's [P][groups][nc]
    // block of eta (eta_dot and eta_dot_dot follow at nh_stride and 2 nh_stride)
    int nh_nc;
    double nh_Q1, nh_Qi, nh_required, nh_tbeta;
    double* nh;
    unsigned long long nh_stride;
    double* modes;              // Nose-Hoover with normal-mode coupling: this replica's [P][D][N] modes
};

struct BatchPtrs {
    double *x, *p, *f, *V;
    const double *Ct, *CTt;
    const RepDesc* reps;
    const int* sched;          // replicas of workgroup g: sched[sched_off[g] .. sched_off[g+1])
    const int* sched_off;
    int* status;               // per replica: 1 when the exchange overflowed
    unsigned long long* ctl;   // [0] step counter, [1] finished-workgroup counter
    double* cmp;               // per replica [8]: pending COM momentum [0..3], known sum of p [4..7]
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

// Wave-uniform copies (v_readfirstlane) of values that reach a non-inlined
// function in vector registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

// Wave-uniform copy of a replica descriptor.  The phases below are kept out
// of line (each one's registers would otherwise add to the others'); their
// arguments arrive in vector registers, and these copies put the replica's
// sizes, constants and offsets back into scalar registers.  (The fields
// of the normal-mode propagator and the Nose-Hoover chains are read where
// they are used: the copy is kept as the other phases need it.)
__device__ __forceinline__ RepDesc rep_uniform(const RepDesc& r) {
    RepDesc u;
    u.N = uni(r.N); u.P = uni(r.P); u.bosonic = uni(r.bosonic); u.pbc = uni(r.pbc);
    u.thermostat = uni(r.thermostat); u.nm = uni(r.nm); u.fixcom = uni(r.fixcom);
    u.plds = uni(r.plds); u.fused = uni(r.fused); u.xslots = uni(r.xslots);
    u.L = uni(r.L); u.invL = uni(r.invL); u.mass = uni(r.mass); u.dt = uni(r.dt);
    u.k = uni(r.k); u.bt = uni(r.bt); u.c1 = uni(r.c1); u.c2 = uni(r.c2);
    u.int_k = uni(r.int_k); u.strength = uni(r.strength); u.cutoff = uni(r.cutoff); u.cut2 = uni(r.cut2);
    u.ext.kind = uni(r.ext.kind); u.ext.a = uni(r.ext.a); u.ext.b = uni(r.ext.b); u.ext.c = uni(r.ext.c);
    u.ext.d = uni(r.ext.d);
    u.seed = (uint32_t)uni((int)r.seed); u.rid = (uint32_t)uni((int)r.rid);
    u.off_x = (unsigned long long)uni((double*)r.off_x);
    u.off_V = (unsigned long long)uni((double*)r.off_V);
    u.off_C = (unsigned long long)uni((double*)r.off_C);
    return u;
}

// Read-only matrices through the constant address space: wave-uniform reads
// of it become scalar loads.
typedef const __attribute__((address_space(4))) double* const_dptr;

// Modes (and beads) per wave in the normal-mode transforms; even, so that the
// two modes of a Box-Muller pair land on the same lane.
constexpr int NM_BLOCK = 4;

// Normal-mode transforms on the matrix cores (v_mfma_f64_16x16x4f64) for
// P <= NM_MFMA_MAXP.  Lane layout of that instruction (lane l, item v):
// A[l%16][l/16], B[l/16][l%16], D[l/16 + 4v][l%16].
constexpr int NM_MFMA_MAXP = 64;
typedef double v4d __attribute__((ext_vector_type(4)));

__host__ __device__ inline int nm_pad16(int P) { return (P + 15) & ~15; }
// Column blocks of 16 that one round of the transform covers: one wave per
// (16-mode block, column block).
__host__ __device__ inline int nm_mfma_groups(int P, int nwaves) {
    const int mb = nm_pad16(P) / 16;
    return nwaves >= mb ? nwaves / mb : 1;
}
// LDS doubles of the matrix-core path: C^T padded to Pp x Pp and two noise
// tiles [groups][Pp][16].
__host__ __device__ inline size_t nm_mfma_lds_doubles(int P, int nwaves) {
    const size_t Pp = (size_t)nm_pad16(P);
    return Pp * Pp + 2 * (size_t)nm_mfma_groups(P, nwaves) * Pp * 16;
}

// Element loops over a replica's [bead][axis][particle] arrays take
// STREAM_UNROLL elements per thread per pass and issue all their loads before
// any store, so each thread keeps that many loads in flight.
constexpr int STREAM_UNROLL = 4;
// Columns per LDS tile of the normal-mode transforms of wg_nm_drift (LDS:
// 4 P NMD_TC doubles) and wg_nh_ostep (P NMD_TC).
constexpr int NMD_TC = 16;

// Axis of element e of a [bead][axis][particle] array.
template <int D>
__device__ __forceinline__ int elem_axis(int e, int N) {
    return D == 1 ? 0 : (e / N) % D;
}

// p[b][a][*] -= cm[a] for every bead and axis of a replica.
template <int D>
__device__ __forceinline__ void wg_sub_com(double* p, int N, int P, const double (&cm)[D]) {
    const int n = P * D * N, nthr = blockDim.x;
    for (int base = threadIdx.x; base < n; base += STREAM_UNROLL * nthr) {
        double pv[STREAM_UNROLL];
#pragma unroll
        for (int u = 0; u < STREAM_UNROLL; ++u) {
            const int e = base + u * nthr;
            if (e < n) pv[u] = p[e];
        }
#pragma unroll
        for (int u = 0; u < STREAM_UNROLL; ++u) {
            const int e = base + u * nthr;
            if (e < n) p[e] = pv[u] - cm[elem_axis<D>(e, N)];
        }
    }
}

// Optional half kick, O-step (Cartesian or normal-mode coupling) and the
// centre-of-mass momentum of the result, for one replica (langevin.cpp:15-27,
// thermostat_coupling.cpp, simulation.cpp:581-603).  With fixcom the COM
// momentum per axis is returned in cm and, when `subtract`, removed from p
// here; otherwise the caller removes it (wg_kick_drift fuses it).
//
// Replicas with r.fused (normal-mode Langevin on the matrix cores) go
// further.  The second O-step of a step leaves its COM momentum pending in
// cmp[0..D) instead of a pass over p; the first O-step of the next step
// subtracts it from p on reading.  The first O-step also does B + A in its
// epilogue (p -= COM, p += dt/2 f, x += dt p / m; wg_kick_drift's
// arithmetic), with the COM of its result known before the pass: the sum of
// C^T xi over the beads of a column is sqrt(P) xi_0 (only the centroid row of
// C has a nonzero sum), so the COM is (c1 S + c2 sqrt(P) sum_i xi_0) / (N P),
// S the sum of p before the step (cmp[4..]: 0 after a deferred subtraction,
// the sum of the state a caller set).  p then reads back through
// pimdb_batch_get_state with the pending COM removed.
// EXT: the kernel of a batch with normal-mode-propagator or Nose-Hoover
// replicas (the former's kicks take fx).
template <int D, bool EXT = false>
__device__ __noinline__ void wg_momenta(const RepDesc& r_, const BatchPtrs& bp, bool kick_, uint32_t step_,
                                        uint32_t ostep_, double* red_, double* lds_, bool subtract_, double (&cm)[D],
                                        double* cmp_) {
    const RepDesc r = rep_uniform(r_);
    const bool kick = uni((int)kick_) != 0, subtract = uni((int)subtract_) != 0;
    double* cmp = uni(cmp_);
    const bool drift = r.fused && !kick;
    const uint32_t step = (uint32_t)uni((int)step_), ostep = (uint32_t)uni((int)ostep_);
    double* red = uni(red_);
    double* lds = uni(lds_);
    const int N = r.N, P = r.P;
    double* p = uni(bp.p) + r.off_x;
    // the kick's forces: the physical ones alone with the normal-mode propagator
    const double* f = (EXT && uni(r_.nmprop)) ? uni((const double*)r_.fx) : uni(bp.f) + r.off_x;
    const double half_dt = 0.5 * r.dt;
    const int ncols = D * N;
    double csum[D];   // this thread's share of the momentum sums per axis
    double fsum[D];   // fused first O-step: the sum of f per axis
#pragma unroll
    for (int a = 0; a < D; ++a) {
        csum[a] = 0.0;
        fsum[a] = 0.0;
    }
    if (r.thermostat == 1 && r.nm && P <= NM_MFMA_MAXP) {
        // Matrix-core form.  One friction serves every mode (langevin.cpp:15-27,
        // thermostat_coupling.cpp:27-47), and C is orthogonal, so the O-step
        //     p' = C^T (c1 C p + c2 xi) = c1 p + C^T (c2 xi)
        // transforms only the noise.  xi, drawn per mode with the keys of the
        // single-system form (modes 2m and 2m+1 from one Box-Muller pair), goes
        // to LDS as pn[g][k][c] (g = column block of the round); wave (bb, g)
        // multiplies 16 beads of C^T by it on the matrix cores and adds
        // c1 (p + dt/2 f) in the epilogue.  The same p' as the two transforms
        // up to rounding.  LDS: CTl[k][b] = CT[b][k] (Pp x Pp, zero padded).
        const int lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
        const int wave = uni((int)(threadIdx.x / WAVE));
        const int Pp = nm_pad16(P), MB = Pp / 16, KK = Pp / 4;
        const double dt_m = r.dt / r.mass;   // the drift x += (dt / m) p, one division per replica
        const int G = nm_mfma_groups(P, nw), ntask = MB * G;
        const int NCB = (ncols + 15) / 16;
        lds_double* CTl = (lds_double*)lds;
        lds_double* pn = CTl + (size_t)Pp * Pp;
        const double* CTt = uni(bp.CTt) + r.off_C;
        for (int e = threadIdx.x; e < Pp * Pp; e += blockDim.x) {
            const int a = e / Pp, b = e % Pp;
            CTl[e] = (a < P && b < P) ? CTt[a * P + b] : 0.0;
        }
        // first O-step of a fused replica.  When p is pending (cmp[7]: it
        // holds the last step's p after its first B, the second B and O-step
        // still owed), both O-steps run here in one product:
        //   p2 = c1 (p + dt/2 f) + C^T c2 xi2 - cm2,   p1 = c1 p2 + C^T c2 xi1
        //      = c1 (c1 (p + dt/2 f) - cm2) + C^T c2 (c1 xi2 + xi1),
        // xi2 the last step's second O-step noise (keys: this step, O-step 1),
        // xi1 this step's first.  COMs as above, with the sum of p + dt/2 f
        // from the known sum of p (cmp[4..]) and the sum of f.
        double cmin[D], cm1[D], cm2[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            cmin[a] = 0.0;
            cm1[a] = 0.0;
            cm2[a] = 0.0;
        }
        const bool pend = drift && uni(cmp[7]) != 0.0;
        if (drift && r.fixcom) {
            double s1[D], s2[D], sf[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                s1[a] = 0.0;
                s2[a] = 0.0;
                sf[a] = 0.0;
            }
            for (int col = threadIdx.x; col < ncols; col += blockDim.x) {
                double g0, g1, h0 = 0.0, h1;
                gauss2<true>((uint32_t)col, step, ostep, r.rid, r.seed, g0, g1);   // mode 0 of column col
                if (pend) gauss2<true>((uint32_t)col, step, 1u, r.rid, r.seed, h0, h1);
                const int ax = col / N;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    if (a == ax) {
                        s1[a] += g0;
                        s2[a] += h0;
                    }
                }
            }
            // without an external potential every force term (pairs, springs,
            // exchange) is translation invariant: f sums to zero up to rounding
            const int nel = r.ext.kind != POT_FREE ? P * ncols : 0;
            for (int e = threadIdx.x; e < nel; e += blockDim.x) {
                const double fe = f[e];
                const int ax = elem_axis<D>(e, N);
#pragma unroll
                for (int a = 0; a < D; ++a)
                    if (a == ax) sf[a] += fe;
            }
            const double sq = sqrt((double)P), inv = 1.0 / (double)(N * P);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const double S1 = wg_sum(s1[a], red);
                fsum[a] = wg_sum(sf[a], red);
                if (pend) {
                    const double S2 = wg_sum(s2[a], red);
                    cm2[a] = (r.c1 * (uni(cmp[4 + a]) + half_dt * fsum[a]) + r.c2 * sq * S2) * inv;
                    cm1[a] = r.c2 * sq * S1 * inv;   // p2 sums to zero
                } else {
                    cmin[a] = uni(cmp[a]);
                    cm1[a] = (r.c1 * uni(cmp[4 + a]) + r.c2 * sq * S1) * inv;
                }
            }
        }
        double* x = uni(bp.x) + r.off_x;
        const int q = lane / 16, c16 = lane % 16;
        const int half = (P + 1) / 2, width = G * 16, npair = (Pp / 2) * width;
        // c2 xi of the round starting at column block cb into buf (every thread its share)
        // pair e = m width + gc of a round (modes 2m, 2m+1 of column gc)
        auto pair_noise = [&](int m, int gc, int c0n, lds_double* buf) {
            const int col = c0n + gc;
            double g0 = 0.0, g1 = 0.0;
            if (m < half && col < ncols) {
                gauss2<true>((uint32_t)((size_t)m * ncols + col), step, ostep, r.rid, r.seed, g0, g1);
                if (pend) {   // c1 xi2 + xi1
                    double h0, h1;
                    gauss2<true>((uint32_t)((size_t)m * ncols + col), step, 1u, r.rid, r.seed, h0, h1);
                    g0 = fma(r.c1, h0, g0);
                    g1 = fma(r.c1, h1, g1);
                }
            }
            lds_double* out = buf + (size_t)(gc / 16) * Pp * 16 + gc % 16;
            out[(2 * m) * 16] = 2 * m < P ? r.c2 * g0 : 0.0;
            out[(2 * m + 1) * 16] = 2 * m + 1 < P ? r.c2 * g1 : 0.0;
        };
        // When the round's width divides the workgroup (16 G columns, G =
        // waves / mode blocks: the usual case), a thread keeps one column and
        // steps m by blockDim / width, the same pairs in the same order
        // without a division per pair.
        const bool colfixed = blockDim.x % width == 0;
        const int gcf = threadIdx.x % width, m0f = threadIdx.x / width, mstep = blockDim.x / width;
        auto gen_noise = [&](int cb, lds_double* buf) {
            const int c0n = cb * 16;
            if (colfixed) {
                for (int m = m0f; m < Pp / 2; m += mstep) pair_noise(m, gcf, c0n, buf);
            } else {
                for (int e = threadIdx.x; e < npair; e += blockDim.x) pair_noise(e / width, e % width, c0n, buf);
            }
        };
        // two noise tiles: a round's products read one while the next round's
        // noise is drawn into the other, between the epilogue's loads and its
        // arithmetic, so the draws overlap the memory traffic
        lds_double* const pn0 = pn;
        lds_double* const pn1 = pn + (size_t)G * Pp * 16;
        gen_noise(0, pn0);
        __syncthreads();
        for (int cb0 = 0, rnd = 0; cb0 < NCB; cb0 += G, ++rnd) {
            const int c0 = cb0 * 16;
            const lds_double* cur = (rnd & 1) ? pn1 : pn0;
            lds_double* nxt = (rnd & 1) ? pn0 : pn1;
            const bool more = cb0 + G < NCB;
            bool drawn = false;
            for (int task = wave; task < ntask; task += nw) {
                const int bb = task % MB, g = task / MB;
                const int col = c0 + g * 16 + c16;
                const bool live = cb0 + g < NCB;   // wave-uniform
                v4d acc = {0.0, 0.0, 0.0, 0.0};
                double pv[4], fv[4], xv[4];
                if (live) {
                    const lds_double* B = cur + (size_t)g * Pp * 16;
                    for (int kk = 0; kk < KK; ++kk) {
                        const int k = 4 * kk + q;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(CTl[(size_t)k * Pp + bb * 16 + c16],
                                                                    B[k * 16 + c16], acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int b = bb * 16 + q + 4 * v;
                        pv[v] = 0.0;
                        fv[v] = 0.0;
                        xv[v] = 0.0;
                        if (b < P && col < ncols) {
                            const size_t e = (size_t)b * ncols + col;
                            pv[v] = p[e];
                            if (kick || drift) fv[v] = f[e];
                            if (drift) xv[v] = x[e];
                        }
                    }
                }
                if (!drawn) {
                    if (more) gen_noise(cb0 + G, nxt);
                    drawn = true;
                }
                // acc[v] = (C^T c2 xi) of bead 16bb + q + 4v, column c0 + 16g + c16
                if (live && col < ncols) {
                    const int ax = col / N;
                    double cmi = 0.0, cmo = 0.0, cmt = 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        if (a == ax) {
                            cmi = cmin[a];
                            cmo = cm1[a];
                            cmt = cm2[a];
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int b = bb * 16 + q + 4 * v;
                        if (b < P) {
                            const size_t e = (size_t)b * ncols + col;
                            if (drift) {
                                const double pnew = pend ? r.c1 * (r.c1 * (pv[v] + half_dt * fv[v]) - cmt) + acc[v]
                                                         : r.c1 * (pv[v] - cmi) + acc[v];
                                double w = pnew - cmo;
                                w = w + half_dt * fv[v];
                                p[e] = w;
                                x[e] = fma(dt_m, w, xv[v]);
                            } else {
                                const double pk = kick ? pv[v] + half_dt * fv[v] : pv[v];
                                const double pnew = r.c1 * pk + acc[v];
                                p[e] = pnew;
#pragma unroll
                                for (int a = 0; a < D; ++a)
                                    if (a == ax) csum[a] += pnew;
                            }
                        }
                    }
                }
            }
            if (!drawn && more) gen_noise(cb0 + G, nxt);
            __syncthreads();   // the next round reads nxt; cur is redrawn the round after
        }
    } else if (r.thermostat != 1 || !r.nm) {
        // beads 2m and 2m+1 of a column share one Box-Muller pair
        const int half = (P + 1) / 2, n2 = half * ncols, nthr = blockDim.x;
        constexpr int U = STREAM_UNROLL / 2;
        for (int base = threadIdx.x; base < n2; base += U * nthr) {
            double pv[U][2], fv[U][2];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = base + u * nthr, m = e2 / ncols, col = e2 % ncols;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int b = 2 * m + h;
                    pv[u][h] = 0.0;
                    fv[u][h] = 0.0;
                    if (e2 < n2 && b < P) {
                        const size_t e = (size_t)b * ncols + col;
                        pv[u][h] = p[e];
                        if (kick) fv[u][h] = f[e];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = base + u * nthr, m = e2 / ncols, col = e2 % ncols;
                if (e2 >= n2) continue;
                const int ax = col / N;
                double g[2] = {0.0, 0.0};
                if (r.thermostat == 1) gauss2<true>((uint32_t)e2, step, ostep, r.rid, r.seed, g[0], g[1]);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int b = 2 * m + h;
                    if (b < P) {
                        double v = pv[u][h];
                        if (kick) v += half_dt * fv[u][h];
                        if (r.thermostat == 1) v = r.c1 * v + r.c2 * g[h];
                        p[(size_t)b * ncols + col] = v;
#pragma unroll
                        for (int a = 0; a < D; ++a)
                            if (a == ax) csum[a] += v;
                    }
                }
            }
        }
    } else {
        // Ct[j*P + k] = C[k][j], CTt[k*P + b] = CT[b][k]; both padded by
        // NM_BLOCK doubles so a block may read past the last row.
        const const_dptr Ct = (const_dptr)(uni(bp.Ct) + r.off_C);
        const const_dptr CTt = (const_dptr)(uni(bp.CTt) + r.off_C);
        // Columns (a, i) in tiles of 64 staged in LDS (after the half kick),
        // lane = column.  Wave w computes modes (then beads) 4w .. 4w+3,
        // 4(w+nw) .. of its column, so one LDS read feeds four FMAs and the
        // matrix entries are the same for the whole wave.
        const int lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
        double* pc = lds;                        // [P][64]
        double* pn = lds + (size_t)P * WAVE;     // [P][64]
        for (int c0 = 0; c0 < ncols; c0 += WAVE) {
            const int col = c0 + lane;
            const bool valid = col < ncols;
            const int ax = col / N;
            for (int e = threadIdx.x; e < P * WAVE; e += blockDim.x) {
                const int j = e / WAVE, c = c0 + e % WAVE;
                double v = 0.0;
                if (c < ncols) {
                    const size_t g = (size_t)j * ncols + c;
                    v = p[g];
                    if (kick) v += half_dt * f[g];
                }
                pc[e] = v;
            }
            __syncthreads();
            // modes: sum_j C[k][j] p_j, then the Langevin update; modes 2m
            // and 2m+1 share one Box-Muller pair
            for (int kb = NM_BLOCK * wave; kb < P; kb += NM_BLOCK * nw) {
                double sm[NM_BLOCK];
#pragma unroll
                for (int q = 0; q < NM_BLOCK; ++q) sm[q] = 0.0;
#pragma unroll 4
                for (int j = 0; j < P; ++j) {
                    const double v = pc[j * WAVE + lane];
                    const const_dptr cr = Ct + (size_t)j * P + kb;
#pragma unroll
                    for (int q = 0; q < NM_BLOCK; ++q) sm[q] += cr[q] * v;
                }
#pragma unroll
                for (int h = 0; h < NM_BLOCK; h += 2) {
                    const int k = kb + h;
                    if (k < P) {
                        double g0 = 0.0, g1 = 0.0;
                        if (valid) gauss2<true>((uint32_t)((size_t)(k >> 1) * ncols + col), step, ostep, r.rid, r.seed, g0, g1);
                        pn[k * WAVE + lane] = r.c1 * sm[h] + r.c2 * g0;
                        if (k + 1 < P) pn[(k + 1) * WAVE + lane] = r.c1 * sm[h + 1] + r.c2 * g1;
                    }
                }
            }
            __syncthreads();
            for (int bb = NM_BLOCK * wave; bb < P; bb += NM_BLOCK * nw) {
                double sm[NM_BLOCK];
#pragma unroll
                for (int q = 0; q < NM_BLOCK; ++q) sm[q] = 0.0;
#pragma unroll 4
                for (int k = 0; k < P; ++k) {
                    const double v = pn[k * WAVE + lane];
                    const const_dptr cr = CTt + (size_t)k * P + bb;
#pragma unroll
                    for (int q = 0; q < NM_BLOCK; ++q) sm[q] += cr[q] * v;
                }
                if (valid) {
#pragma unroll
                    for (int q = 0; q < NM_BLOCK; ++q) {
                        if (bb + q < P) {
                            p[(size_t)(bb + q) * ncols + col] = sm[q];
#pragma unroll
                            for (int a = 0; a < D; ++a)
                                if (a == ax) csum[a] += sm[q];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (drift) {
        // p now holds this step's p after its first B: the second B and
        // O-step are owed (done by the next step's first O-step, or by
        // pimdb_batch_get_state's finishing launch); p sums to dt/2 sum f
#pragma unroll
        for (int a = 0; a < D; ++a) cm[a] = 0.0;   // removed in the epilogue
        __syncthreads();   // every thread has read cmp
        if (threadIdx.x == 0) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                cmp[a] = 0.0;
                cmp[4 + a] = half_dt * fsum[a];
            }
            cmp[7] = 1.0;
        }
    } else if (r.fixcom) {
#pragma unroll
        for (int a = 0; a < D; ++a) cm[a] = wg_sum(csum[a], red) / (double)(N * P);
        if (r.fused) {   // left pending for the next step's first O-step
            if (threadIdx.x == 0) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    cmp[a] = cm[a];
                    cmp[4 + a] = 0.0;
                }
                cmp[7] = 0.0;
            }
        } else if (subtract) {
            __syncthreads();
            wg_sub_com<D>(p, N, P, cm);
        }
    } else {
#pragma unroll
        for (int a = 0; a < D; ++a) cm[a] = 0.0;
        if (r.fused && threadIdx.x == 0) cmp[7] = 0.0;   // the owed O-step is done
    }
    __syncthreads();
}

// One Nose-Hoover O-step of a replica (nose_hoover.cpp:96-150 per chain, as
// the single-system k_nh_ostep): the optional half kick (p += dt/2 f), then
// for every bead (Cartesian coupling) or every normal mode (normal-mode
// coupling: the modes C p of the kicked momenta, with the single-system
// k_momenta stage 1's arithmetic, and back with C^T after, stage 2's) the
// chains, one per bead, per particle or per degree of freedom, each advanced
// on the kinetic measure of its own degrees of freedom (nh_chain_step), and
// the scaling; then the centre-of-mass momentum of the result in cm (fixcom),
// removed here when `subtract` (the single-system form removes it at the
// start of the next O-step; nothing reads p in between).  Waves take beads
// (modes); the transforms take column tiles of NMD_TC in LDS.
template <int D, bool EXT>
__device__ __noinline__ void wg_nh_ostep(const RepDesc& r_, const BatchPtrs& bp, bool kick_, bool subtract_,
                                         double* red_, double* lds_, double (&cm)[D]) {
    const RepDesc r = rep_uniform(r_);
    const bool kick = uni((int)kick_) != 0, subtract = uni((int)subtract_) != 0;
    double* red = uni(red_);
    const int N = r.N, P = r.P, kind = r.thermostat;
    const int lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
    const int wave = uni((int)(threadIdx.x / WAVE));
    double* p = uni(bp.p) + r.off_x;
    const double* f = (EXT && uni(r_.nmprop)) ? uni((const double*)r_.fx) : uni(bp.f) + r.off_x;
    NhArgs h{};
    h.nc = uni(r_.nh_nc);
    h.dt = r.dt;
    h.Q1 = uni(r_.nh_Q1); h.Qi = uni(r_.nh_Qi); h.required = uni(r_.nh_required); h.thermo_beta = uni(r_.nh_tbeta);
    const int groups = kind == PIMDB_THERMO_NOSE_HOOVER ? 1 : kind == PIMDB_THERMO_NOSE_HOOVER_NP ? N : N * D;
    const unsigned long long nh_stride = (unsigned long long)uni((double*)r_.nh_stride);
    double* eta = uni(r_.nh);
    double* eta_dot = eta + nh_stride;
    double* eta_dot_dot = eta_dot + nh_stride;
    const double half_dt = 0.5 * r.dt;
    const bool nmc = r.nm != 0;   // normal-mode coupling
    const int ncols = D * N;
    double* lds = uni(lds_);
    const double* Ct = uni(bp.Ct) + r.off_C;     // Ct[j*P + k] = C[k][j]
    const double* CTt = uni(bp.CTt) + r.off_C;   // CTt[k*P + b] = CT[b][k]
    double* modes = nmc ? uni(r_.modes) : nullptr;
    double csum[D];
#pragma unroll
    for (int a = 0; a < D; ++a) csum[a] = 0.0;
    if (nmc) {
        // the modes of the kicked momenta
        double* pc = lds;   // [P][NMD_TC]
        for (int c0 = 0; c0 < ncols; c0 += NMD_TC) {
            const int nc = min(NMD_TC, ncols - c0);
            for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
                const int b = e / NMD_TC, c = e % NMD_TC;
                double v = 0.0;
                if (c < nc) {
                    const size_t g = (size_t)b * ncols + c0 + c;
                    v = p[g];
                    if (kick) v += half_dt * f[g];
                }
                pc[e] = v;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
                const int k = e / NMD_TC, c = e % NMD_TC;
                if (c < nc) {
                    double sm = 0.0;
                    for (int j = 0; j < P; ++j) sm += Ct[(size_t)j * P + k] * pc[j * NMD_TC + c];
                    modes[(size_t)k * ncols + c0 + c] = sm;
                }
            }
            __syncthreads();
        }
    }
    double* q = nmc ? modes : p;   // the momenta the chains scale: per bead, or per mode
    for (int b = wave; b < P; b += nw) {
        double* pb = q + (size_t)b * D * N;
        const size_t cbase = (size_t)b * groups * h.nc;
        if (kick && !nmc) {
            const double* fb = f + (size_t)b * D * N;
            for (int e = lane; e < D * N; e += WAVE) pb[e] = pb[e] + half_dt * fb[e];
            __builtin_amdgcn_wave_barrier();
        }
        if (kind == PIMDB_THERMO_NOSE_HOOVER) {
            double s = 0.0;
            for (int e = lane; e < D * N; e += WAVE) s += pb[e] * pb[e];
            s = wave_sum(s);
            double sc = 0.0;
            if (lane == 0) sc = nh_chain_step(h, eta + cbase, eta_dot + cbase, eta_dot_dot + cbase, s / r.mass);
            sc = wave_bcast_uniform(sc, 0);
            for (int e = lane; e < D * N; e += WAVE) pb[e] = pb[e] * sc;
        } else if (kind == PIMDB_THERMO_NOSE_HOOVER_NP) {
            for (int i = lane; i < N; i += WAVE) {
                double ce = 0.0;
#pragma unroll
                for (int a = 0; a < D; ++a) ce += pb[a * N + i] * pb[a * N + i];
                ce /= r.mass;
                const size_t o = cbase + (size_t)i * h.nc;
                const double sc = nh_chain_step(h, eta + o, eta_dot + o, eta_dot_dot + o, ce);
#pragma unroll
                for (int a = 0; a < D; ++a) pb[a * N + i] = pb[a * N + i] * sc;
            }
        } else {
            for (int e = lane; e < D * N; e += WAVE) {
                const int a = e / N, i = e % N;
                const double ce = pb[e] * pb[e] / r.mass;
                const size_t o = cbase + ((size_t)i * D + a) * h.nc;
                const double sc = nh_chain_step(h, eta + o, eta_dot + o, eta_dot_dot + o, ce);
                pb[e] = pb[e] * sc;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (!nmc) {
            for (int e = lane; e < D * N; e += WAVE) {
                const double v = pb[e];
#pragma unroll
                for (int a = 0; a < D; ++a)
                    if (a == e / N) csum[a] += v;
            }
        }
    }
    __syncthreads();
    if (nmc) {
        // back to Cartesian momenta
        double* pn = lds;   // [P][NMD_TC]
        for (int c0 = 0; c0 < ncols; c0 += NMD_TC) {
            const int nc = min(NMD_TC, ncols - c0);
            for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
                const int k = e / NMD_TC, c = e % NMD_TC;
                pn[e] = c < nc ? modes[(size_t)k * ncols + c0 + c] : 0.0;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
                const int b = e / NMD_TC, c = e % NMD_TC;
                if (c < nc) {
                    double sm = 0.0;
                    for (int k = 0; k < P; ++k) sm += CTt[(size_t)k * P + b] * pn[k * NMD_TC + c];
                    p[(size_t)b * ncols + c0 + c] = sm;
#pragma unroll
                    for (int a = 0; a < D; ++a)
                        if (a == (c0 + c) / N) csum[a] += sm;
                }
            }
            __syncthreads();
        }
    }
    if (r.fixcom) {
#pragma unroll
        for (int a = 0; a < D; ++a) cm[a] = wg_sum(csum[a], red) / (double)(N * P);
        if (subtract) {
            __syncthreads();
            wg_sub_com<D>(p, N, P, cm);
        }
    } else {
#pragma unroll
        for (int a = 0; a < D; ++a) cm[a] = 0.0;
    }
    __syncthreads();
}

// zeroMomentum (cm, when fixcom), then B + A: p += dt/2 f, x += dt p / m
// (simulation.cpp:581-603, velocity_verlet.cpp:24-38).
template <int D>
__device__ __noinline__ void wg_kick_drift(const RepDesc& r_, const BatchPtrs& bp, const double (&cm_)[D]) {
    const RepDesc r = rep_uniform(r_);
    double cm[D];
#pragma unroll
    for (int a = 0; a < D; ++a) cm[a] = uni(cm_[a]);
    const int N = r.N, P = r.P;
    double* x = uni(bp.x) + r.off_x;
    double* p = uni(bp.p) + r.off_x;
    const double* f = uni(bp.f) + r.off_x;
    const double half_dt = 0.5 * r.dt;
    const double dt_m = r.dt / r.mass;   // x += (dt / m) p: one division per replica
    const int n = P * D * N, nthr = blockDim.x;
    for (int base = threadIdx.x; base < n; base += STREAM_UNROLL * nthr) {
        double pv[STREAM_UNROLL], fv[STREAM_UNROLL], xv[STREAM_UNROLL];
#pragma unroll
        for (int u = 0; u < STREAM_UNROLL; ++u) {
            const int e = base + u * nthr;
            if (e < n) {
                pv[u] = p[e];
                fv[u] = f[e];
                xv[u] = x[e];
            }
        }
#pragma unroll
        for (int u = 0; u < STREAM_UNROLL; ++u) {
            const int e = base + u * nthr;
            if (e < n) {
                double v = pv[u];
                if (r.fixcom) v -= cm[elem_axis<D>(e, N)];
                v = v + half_dt * fv[u];
                p[e] = v;
                x[e] = fma(dt_m, v, xv[u]);
            }
        }
    }
    __syncthreads();
}

// The normal-mode propagator's drift for a replica of distinguishable
// particles (normal_modes_propagator.cpp:19-60; the reference allows it only
// without exchange): zeroMomentum (cm, when fixcom) and B (p += dt/2 fx, the
// physical forces), then the free ring polymer over dt in normal
// modes: x_nm = C x and p_nm = C p, each mode rotated by its harmonic
// propagator (the centroid, w = 0, moves freely), and back with C^T.
// Columns (axis, particle) in tiles of NMD_TC staged in LDS; the arithmetic
// is the single-system form's (k_nm_propagate), element for element.
template <int D>
__device__ __noinline__ void wg_nm_drift(const RepDesc& r_, const BatchPtrs& bp, const double (&cm_)[D], double* lds_) {
    const RepDesc r = rep_uniform(r_);
    double cm[D];
#pragma unroll
    for (int a = 0; a < D; ++a) cm[a] = uni(cm_[a]);
    double* lds = uni(lds_);
    const int N = r.N, P = r.P, ncols = D * N;
    double* x = uni(bp.x) + r.off_x;
    double* p = uni(bp.p) + r.off_x;
    const double* f = uni((const double*)r_.fx);   // physical forces
    const double* Ct = uni(bp.Ct) + r.off_C;     // Ct[j*P + k] = C[k][j]
    const double* CTt = uni(bp.CTt) + r.off_C;   // CTt[k*P + b] = CT[b][k]
    const double* rot = uni(r_.rot);
    const double half_dt = 0.5 * r.dt;
    double* xc = lds;                          // [P][NMD_TC] Cartesian x, p of the tile's columns
    double* pc = xc + (size_t)P * NMD_TC;
    double* qn = pc + (size_t)P * NMD_TC;      // [P][NMD_TC] propagated modes
    double* pn = qn + (size_t)P * NMD_TC;
    for (int c0 = 0; c0 < ncols; c0 += NMD_TC) {
        const int nc = min(NMD_TC, ncols - c0);
        for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
            const int b = e / NMD_TC, c = e % NMD_TC;
            double xv = 0.0, pv = 0.0;
            if (c < nc) {
                const size_t g = (size_t)b * ncols + c0 + c;
                xv = x[g];
                pv = p[g];
                if (r.fixcom) pv -= cm[(c0 + c) / N];
                pv += half_dt * f[g];
            }
            xc[e] = xv;
            pc[e] = pv;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
            const int k = e / NMD_TC, c = e % NMD_TC;
            double q = 0.0, pp = 0.0;
            for (int j = 0; j < P; ++j) {
                const double cj = Ct[(size_t)j * P + k];
                q += cj * xc[j * NMD_TC + c];
                pp += cj * pc[j * NMD_TC + c];
            }
            const double cs = rot[4 * k], sn = rot[4 * k + 1], mw = rot[4 * k + 2], w = rot[4 * k + 3];
            if (w == 0) {
                qn[e] = q + r.dt / r.mass * pp;
                pn[e] = pp;
            } else {
                qn[e] = cs * q + sn / mw * pp;
                pn[e] = (-1) * mw * sn * q + cs * pp;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < P * NMD_TC; e += blockDim.x) {
            const int b = e / NMD_TC, c = e % NMD_TC;
            if (c >= nc) continue;
            double xv = 0.0, pv = 0.0;
            for (int k = 0; k < P; ++k) {
                const double ck = CTt[(size_t)k * P + b];
                xv += ck * qn[k * NMD_TC + c];
                pv += ck * pn[k * NMD_TC + c];
            }
            const size_t g = (size_t)b * ncols + c0 + c;
            x[g] = xv;
            p[g] = pv;
        }
        __syncthreads();
    }
}

// Physical forces and interior springs on every (bead, particle) of a replica
// (simulation.cpp:353-455).  Each wave owns whole beads: the pairs of a bead
// with Newton's third law (slice_pair_forces), then the trap and the springs.
// after_exch_: the exchange phase has already put its forces into f of beads
// 1 and P (only with the pair forces in LDS), which then get the physical
// forces added instead of written.
// NMP: replicas with the normal-mode propagator, whose physical forces also
// go to fx (a separate instantiation, so the other replicas' code is unchanged).
template <int D, int KIND, bool NMP = false>
__device__ __noinline__ void wg_bead_forces(const RepDesc& r_, const BatchPtrs& bp, double* lds_, int after_exch_) {
    const RepDesc r = rep_uniform(r_);
    const bool after_exch = uni(after_exch_) != 0;
    lds_double* F = (lds_double*)uni(lds_);
    const int N = r.N, P = r.P;
    const double* x = uni(bp.x) + r.off_x;
    double* f = uni(bp.f) + r.off_x;
    const bool pbc = r.pbc != 0;
    const double invL_s = pbc ? 1.0 / r.L : 0.0;   // the springs' min image
    const size_t bs = (size_t)D * N;
    const int lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const PairParams pp = make_pair_params(r.L, r.pbc, r.cutoff, r.cut2, r.int_k, r.strength);
    // the fractional-coordinate constants, wave-uniform (scalar registers)
    FracParams fp = make_frac_params(pp);
    fp.cut2 = uni(fp.cut2); fp.kh = uni(fp.kh); fp.kd = uni(fp.kd); fp.ex = uni(fp.ex);
    fp.b0 = uni(fp.b0); fp.b1 = uni(fp.b1); fp.b2 = uni(fp.b2); fp.dl = uni(fp.dl);
    // b2 and b1 meet in one FMA of the pair loop: copies in vector registers
    fp.b1v = vgpr_copy(fp.b1);
    fp.b2v = vgpr_copy(fp.b2);
    // this wave's slice forces by tile position (r.plds), then the tile
    // order of the particles and its inverse
    const size_t fs = (size_t)pair_fstride(D) * N;
    lds_double* Fw = F + (size_t)wave * fs;
    lds_int* perm = (lds_int*)(F + (size_t)nw * fs);   // [N] particle at each tile position
    lds_int* iperm = perm + N;                          // [N] tile position of each particle
    if (KIND != POT_FREE && r.cutoff != 0.0 && r.plds) {
        // Aziz with PBC: a spatial order from bead 0 (pair_order_key), which
        // every bead shares; otherwise the identity.  The keys sit in the
        // force arrays until the slices zero them.
        lds_uint* key = (lds_uint*)F;
        const bool order = KIND == POT_AZIZ && pbc;
        const double invL = 1.0 / r.L;
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            double xi[D];
#pragma unroll
            for (int a = 0; a < D; ++a) xi[a] = x[a * N + i];
            key[i] = order ? pair_order_key<D>(xi, invL, N, i) : (unsigned)i;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            const unsigned ki = key[i];
            int rank = 0;
            for (int j = 0; j < N; ++j) rank += key[j] < ki ? 1 : 0;
            perm[rank] = i;
            iperm[i] = rank;
        }
        __syncthreads();
    }
    for (int b = wave; b < P; b += nw) {
        const double* xs = x + bs * b;
        double* fo = f + bs * b;
        const bool pairs = KIND != POT_FREE && r.cutoff != 0.0;
        const bool in_lds = pairs && r.plds;
        if (pairs) {
            if (in_lds) {
                switch (pair_image(pp)) {
                case 0: slice_pair_forces_lds<D, KIND, 0>(xs, N, WalkGrad<D, KIND, 0>{pp, fp}, lane, Fw, perm); break;
                case 1: slice_pair_forces_lds<D, KIND, 1>(xs, N, WalkGrad<D, KIND, 1>{pp, fp}, lane, Fw, perm); break;
                default: slice_pair_forces_lds<D, KIND, 2>(xs, N, WalkGrad<D, KIND, 2>{pp, fp}, lane, Fw, perm); break;
                }
            } else {
                switch (pair_image(pp)) {
                case 0: slice_pair_forces<D, KIND, 0>(xs, fo, N, pp, lane); break;
                case 1: slice_pair_forces<D, KIND, 1>(xs, fo, N, pp, lane); break;
                default: slice_pair_forces<D, KIND, 2>(xs, fo, N, pp, lane); break;
                }
            }
        } else {
            for (int e = lane; e < D * N; e += WAVE) fo[e] = 0.0;
            __threadfence_block();
        }
        // trap and interior springs; exchange beads get their spring part
        // from the exchange phase
        const bool exchange_bead = r.bosonic && (b == 0 || b == P - 1);
        const double* xprev = x + bs * ((b - 1 + P) % P);
        const double* xnext = x + bs * ((b + 1) % P);
        for (int i = lane; i < N; i += WAVE) {
            double xv[D], fe[D];
            const int ip = in_lds ? iperm[i] * pair_fstride(D) : 0;
#pragma unroll
            for (int a = 0; a < D; ++a) xv[a] = xs[a * N + i];
            ext_force<D>(r.ext, xv, fe);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const int e = a * N + i;
                const double phys = fe[a] + (in_lds ? Fw[ip + a] : fo[e]);
                double spring = 0.0;
                if (!exchange_bead) {
                    // (the reference's image, min_image, without its division
                    // except next to a tie)
                    const double dp = disp_fast(xv[a], xprev[e], pbc, r.L, invL_s);
                    const double dn = disp_fast(xv[a], xnext[e], pbc, r.L, invL_s);
                    spring = r.k * (dp + dn);
                }
                fo[e] = exchange_bead ? (after_exch ? fo[e] + phys : phys) : spring + phys;
                if constexpr (NMP) uni(r_.fx)[bs * b + e] = phys;
            }
        }
    }
    __syncthreads();
}

// LDS doubles of the integrator's normal-mode transforms (matrix-core form
// for P <= NM_MFMA_MAXP, column tiles of 64 otherwise) with nwaves waves.
__host__ __device__ inline size_t batch_nm_lds_doubles(int P, int nwaves) {
    return P <= NM_MFMA_MAXP ? nm_mfma_lds_doubles(P, nwaves) : 2 * (size_t)P * WAVE;
}

// Exchange of one replica (exchange_batch.hpp): the forward recursion, then
// the backward recursion in linear form with the exterior forces added into f
// of beads 1 and P (add_ = 0: written into them).  Kept out of line (its registers would otherwise add to
// every other phase's); the arguments are made wave-uniform on entry.
// DEV: the workspace lds_ is the replica's device-memory one (RepDesc::xws), a
// separate instantiation so that the LDS one keeps its LDS instructions.  The
// workspace is only touched between workgroup barriers or by the lane that
// wrote it, so the workgroup's ordering of device memory is all it needs.
template <int D, bool DEV = false>
__device__ __noinline__ void wg_exchange(int N_, int P_, int pbc_, double L_, double k_, double bt_, const double* x_,
                                         double* f_, double* V_, int* status_, double* lds_, int slots_, int add_) {
    const int N = uni(N_), P = uni(P_);
    const bool add = uni(add_) != 0;
    const int slots = uni(slots_);
    const double L = uni(L_), k = uni(k_), bt = uni(bt_);
    const double* x = uni(x_);
    double* f = uni(f_);
    double* V = uni(V_);
    int* status = uni(status_);
    double* lds = uni(lds_);
    if (uni(pbc_) != 0) xb_exchange<D, true>(N, P, L, k, bt, x, f, V, status, lds, slots, add);
    else xb_exchange<D, false>(N, P, L, k, bt, x, f, V, status, lds, slots, add);
}

// Work of one launch: a full MD step, or parts of it for timing.
enum BatchMode : int { BM_STEP = 0, BM_FORCES = 1, BM_PAIRS = 2, BM_EXCHANGE = 3, BM_INTEGRATOR = 4, BM_FINISH = 5 };

// One MD step of every replica (BM_STEP), the forces only (BM_FORCES), one
// phase for timing, or the owed second B and O-step of replicas whose p is
// pending (BM_FINISH; see wg_momenta).  The mode is a template parameter so
// that every mode is its own kernel (and its own line in a kernel trace).
// EXT: a batch with normal-mode-propagator or Nose-Hoover replicas, or with
// an exchange workspace in device memory (a separate instantiation, so that
// the kernel of every other batch is unchanged).
template <int D, int KIND, int MODE, bool EXT = false>
static __global__ void __launch_bounds__(1024) k_batch_step(BatchPtrs bp) {
    extern __shared__ double smem[];
    double* red = smem;                 // [16]
    double* lds = smem + 16;            // normal-mode tiles / recursion workspace (not live together)
    const uint32_t step = (uint32_t)bp.ctl[0];
    const int g = blockIdx.x;
    for (int s = bp.sched_off[g]; s < bp.sched_off[g + 1]; ++s) {
        const int ri = bp.sched[s];
        const RepDesc& r = bp.reps[ri];
        constexpr bool integ = MODE == BM_STEP || MODE == BM_INTEGRATOR;
        if constexpr (integ) {
            double cm[D];
            if (EXT && r.thermostat >= PIMDB_THERMO_NOSE_HOOVER)
                wg_nh_ostep<D, EXT>(r, bp, false, false, red, lds, cm);
            else
                wg_momenta<D, EXT>(r, bp, false, step, 0, red, lds, false, cm, bp.cmp + 8 * (size_t)ri);
            if (EXT && r.nmprop) wg_nm_drift<D>(r, bp, cm, lds);
            else if (!r.fused) wg_kick_drift<D>(r, bp, cm);
        }
        // the second half of the grid (with two workgroups per compute unit,
        // mostly the other workgroup of each unit) takes the exchange before
        // the pairs, so that one workgroup's exchange meets the other's pair
        // walk instead of its own exchange
        const bool xfirst = MODE <= BM_FORCES && r.bosonic && r.plds && KIND != POT_FREE && r.cutoff != 0.0 &&
                            2 * (int)blockIdx.x >= (int)gridDim.x;   // (pair forces in LDS: f written once)
        if (xfirst) {
            if (EXT && r.bosonic && r.xws)
                wg_exchange<D, true>(r.N, r.P, r.pbc, r.L, r.k, r.bt, bp.x + r.off_x, bp.f + r.off_x, bp.V + r.off_V,
                                     bp.status + ri, r.xws, r.xslots, 0);
            else
                wg_exchange<D>(r.N, r.P, r.pbc, r.L, r.k, r.bt, bp.x + r.off_x, bp.f + r.off_x, bp.V + r.off_V,
                               bp.status + ri, lds, r.xslots, 0);
        }
        if constexpr (MODE <= BM_PAIRS) {
            if (EXT && r.nmprop) wg_bead_forces<D, KIND, true>(r, bp, lds, 0);   // (no exchange)
            else wg_bead_forces<D, KIND>(r, bp, lds, xfirst ? 1 : 0);
        }
        if ((MODE <= BM_FORCES || MODE == BM_EXCHANGE) && r.bosonic && !xfirst) {
            if (EXT && r.bosonic && r.xws)
                wg_exchange<D, true>(r.N, r.P, r.pbc, r.L, r.k, r.bt, bp.x + r.off_x, bp.f + r.off_x, bp.V + r.off_V,
                                     bp.status + ri, r.xws, r.xslots, 1);
            else
                wg_exchange<D>(r.N, r.P, r.pbc, r.L, r.k, r.bt, bp.x + r.off_x, bp.f + r.off_x, bp.V + r.off_V,
                               bp.status + ri, lds, r.xslots, 1);
        }
        if constexpr (integ) {
            if (!r.fused) {   // fused replicas owe it to the next step's first O-step
                double cm[D];
                if (EXT && r.thermostat >= PIMDB_THERMO_NOSE_HOOVER)
                    wg_nh_ostep<D, EXT>(r, bp, true, true, red, lds, cm);
                else
                    wg_momenta<D, EXT>(r, bp, true, step + 1, 1, red, lds, true, cm, bp.cmp + 8 * (size_t)ri);
            }
        }
        if constexpr (MODE == BM_FINISH) {
            // the O-step keys of the step that left p pending: its step + 1,
            // which is the counter now
            if (r.fused && bp.cmp[8 * (size_t)ri + 7] != 0.0) {
                double cm[D];
                wg_momenta<D>(r, bp, true, step, 1, red, lds, true, cm, bp.cmp + 8 * (size_t)ri);
            }
        }
    }
    if constexpr (MODE == BM_STEP) {
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
