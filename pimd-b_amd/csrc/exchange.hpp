// exchange.hpp -- Feldman-Hirshberg bosonic exchange on the GPU: cycle
// energies, the forward/backward recursions and the exchange forces (the
// adjoint of V_B plus the interior links) on beads 1 and P.
//
// Reference: src/bosonic_exchange/quadratic_bosonic_exchange.cpp and
// bosonic_exchange_base.cpp of higj/pimd-b.
//
// Storage: bead-major SoA positions, x[(b*D + a)*N + i].  The cycle energy of
// the cycle through particles u..v (u <= v) is ecyc(u,v) = E^(v+1, v-u+1) of
// the reference.  On the step path it is recomputed wherever it is needed
// from beads 1 and P and the (split) prefix sums of the ring's links
// (Ring, below: O(N) data, staged in LDS).  Only when a caller asks for the
// energies themselves (or for the observables) is the triangle materialised,
// at E[v(v+1)/2 + u]: exactly the reference's packed E_kn triangle (E(m,k) at
// m(m+1)/2 - k), so column v is contiguous.
//
// Recursions.  Both recursions have the form
//     U_n = -ln sum_{q<n} exp(-(w(n,q) + U_q)),  U_0 = 0,
// forward : U_n = beta~ V[n],      w(n,q) = beta~ ecyc(q, n-1) + ln n
// backward: U_n = beta~ Vb[N-n],   w(n,q) = beta~ ecyc(N-n, N-1-q) + ln(N-q)
// (quadratic_bosonic_exchange.cpp:73-128).  Instead of one log-sum-exp per n
// (a reduction, an exp and a log on the sequential path), targets are taken in
// blocks of REC_B and each block runs
//   (1) one pass over the sources of earlier blocks, giving per target
//       m_n = min_q (w(n,q) + U_q) and s_n = sum_q exp(-(w(n,q) + U_q - m_n))
//       (an online log-sum-exp: the partial sum is rescaled when m drops)
//   (2) a min-plus pass  R_n = min(m_n, min_q (w(n,q) + R_q))  [sequential part: add+min]
//   (3) a linear solve   z_n = s_n exp(-(m_n - R_n)) + sum_q exp(-(w(n,q) + R_q - R_n)) z_q
//                                                         [sequential part: one FMA]
//   (4) U_n = R_n - ln z_n
// where q in (2) and (3) runs over the block's own earlier targets.  Every
// exponent is <= 0 and the minimising term contributes at least 1, so
// 1 <= z_n <= n 2^REC_B: no overflow or underflow for any N.  All
// exponentials are off the sequential path.  The result equals the
// reference's shifted log-sum-exp up to rounding.
#pragma once

#include "device_common.hpp"

namespace pimdb {

constexpr int REC_B = 64;   // targets per block: one per lane of the sequential wave
constexpr int REC_MAXWAVES = 16;
constexpr int REC_UNROLL = 8;  // independent source loads in flight per lane
constexpr int REC_PF = 16;     // in-block entries read ahead by the sequential passes
// The in-block matrix holds its strict lower triangle only, packed by
// target: entry (j, t), j < t, at rec_idx(j, t) = t (t - 1) / 2 + j, so that
// a wave writing one target's sources (the triangle passes) writes
// consecutive doubles.
constexpr int REC_TRI = REC_B * (REC_B - 1) / 2;
__host__ __device__ __forceinline__ int rec_idx(int j, int t) { return t * (t - 1) / 2 + j; }

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

// ---- double-double helpers (error-free transformations; no contraction) ----
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD dd_two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
#pragma clang fp contract(off)
    const DD s = dd_two_sum(a.hi, b.hi);
    const double lo = s.lo + (a.lo + b.lo);
    const double hi = s.hi + lo;
    return DD{hi, lo - (hi - s.hi)};
}

// ---- cycle energies from link prefix sums ----
// With link_j = |x1[j+1] - xP[j]|^2 (min-imaged) and S_v = sum_{j<v} link_j,
//     ecyc(u,v) = k/2 (S_v - S_u + |x1[u] - xP[v]|^2),
// the closed form of the reference's recurrence (quadratic_bosonic_exchange.cpp
// :45-60).  S_v is summed as a double-double and stored split as
// S_v = H_v + L_v, where H_v is S_v rounded to a multiple of the power of two
// q = 2^(e - 40) (2^e <= S_{N-1} < 2^(e+1), the largest sum) and L_v is the
// rest, |L_v| <= q/2.  H_v - H_u is then exact (a multiple of q below 2^42 q)
// and L_v - L_u rounds far below the last bit of the result, so
// (H_v - H_u) + (L_v - L_u) is S_v - S_u rounded once, in three additions:
// any ecyc is a few flops from O(N) data however long the ring, and the
// N(N+1)/2 triangle is never stored on the step path.
struct Ring {
    const double* x1;    // [D][N] bead 1
    const double* xP;    // [D][N] bead P
    const double* Shi;   // [N] H_v, prefix sums rounded to multiples of q
    const double* Slo;   // [N] L_v = S_v - H_v
    double halfk, L, invL;   // L = invL = 0 without periodic boundaries
};

__host__ __device__ inline Ring make_ring(const double* x1, const double* xP, const double* Sh, const double* Sl,
                                          double k, bool pbc, double L) {
    return Ring{x1, xP, Sh, Sl, 0.5 * k, pbc ? L : 0.0, pbc ? 1.0 / L : 0.0};
}

// Minimum image for a squared distance (the image choice at |dx| = L/2 does
// not change the square); one rounding less than the reference's
// dx - L floor(dx/L + 1/2) (bosonic_exchange_base.cpp:30-42).  Without
// periodic boundaries L = invL = 0 and the result is dx exactly.  Odd in dx
// (rint rounds half to even), so the image of -dx is minus the image of dx.
__device__ __forceinline__ double ring_image(double dx, const Ring& r) {
    return fma(-r.L, rint(dx * r.invL), dx);
}

// S_v - S_u from the split sums (see above).
__device__ __forceinline__ double ring_sdiff(double shv, double slv, double shu, double slu) {
#pragma clang fp contract(off)
    return (shv - shu) + (slv - slu);
}

template <int D>
__device__ __forceinline__ double ring_ecyc(const Ring& r, const double (&d)[D], double shu, double slu, double shv,
                                            double slv) {
    return r.halfk * (ring_sdiff(shv, slv, shu, slu) + norm2_fma<D>(d));
}

// S_0..S_{N-1} of the ring by one wave (chunks of ceil(N/64) per lane, then a
// scan of the chunk totals across lanes), stored split as H + L; ring.x1 /
// ring.xP readable, Shi / Slo written (LDS or global).
template <int D>
__device__ __forceinline__ void ring_prefix(const Ring& r, int N, int lane, double* Shi, double* Slo) {
    const int c = (N + WAVE - 1) / WAVE;
    const int v0 = lane * c;
    DD acc{0.0, 0.0};
    for (int t = 0; t < c; ++t) {
        const int v = v0 + t;
        if (v < N) {
            Shi[v] = acc.hi;
            Slo[v] = acc.lo;
            if (v < N - 1) {
                double d[D];
#pragma unroll
                for (int a = 0; a < D; ++a) d[a] = ring_image(r.x1[a * N + v + 1] - r.xP[a * N + v], r);
                acc = dd_add(acc, DD{norm2_fma<D>(d), 0.0});
            }
        }
    }
    DD inc = acc;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const DD o{__shfl_up(inc.hi, off, WAVE), __shfl_up(inc.lo, off, WAVE)};
        if (lane >= off) inc = dd_add(o, inc);
    }
    DD ex{__shfl_up(inc.hi, 1, WAVE), __shfl_up(inc.lo, 1, WAVE)};
    if (lane == 0) ex = DD{0.0, 0.0};
    // q from the total (lane 63's inclusive sum, the largest S)
    const double total = __shfl(inc.hi, WAVE - 1, WAVE);
    const double q = total > 0.0 ? ldexp(1.0, ilogb(total) - 40) : 1.0;
    const double invq = 1.0 / q;
    for (int t = 0; t < c; ++t) {
        const int v = v0 + t;
        if (v < N) {
            const DD sv = dd_add(ex, DD{Shi[v], Slo[v]});
            const double h = rint(sv.hi * invq) * q;
            Shi[v] = h;
            Slo[v] = (sv.hi - h) + sv.lo;
        }
    }
}

// Shared-memory workspace of one workgroup running one recursion.
struct RecShared {
    double* U;         // [N+1]
    double* lg;        // [N+1]  ln(i) (staged with the ring)
    double* R;         // [REC_B] shifts of the block's targets
    double* z;         // [REC_B] the block's forward-substitution results
    double* part;      // [nw][REC_B] per-wave partial min of the earlier sources
    double* psum;      // [nw][REC_B] per-wave partial sums scaled to that min, then to R
    double* a;         // [REC_TRI] in-block weights, then coefficients, at rec_idx(j, t)
    double* x1;        // [D][N] bead 1
    double* xP;        // [D][N] bead P
    double* Shi;       // [N] link prefix sums
    double* Slo;       // [N]
};

// Doubles of the staged ring (beads 1 and P, split prefix sums), of the
// table ln i and of one direction's workspace for a group of nw waves.
__host__ __device__ inline size_t rec_ring_doubles(int N, int D) { return (size_t)(2 * D + 2) * N; }
__host__ __device__ inline size_t rec_lg_doubles(int N) { return (size_t)N + 1; }
__host__ __device__ inline size_t rec_dir_doubles(int N, int nw) {
    return (size_t)(N + 1) + 2 * REC_B + 2 * (size_t)nw * REC_B + (size_t)REC_TRI;
}
// One workgroup of nw waves running one recursion at a time: the ring, ln i,
// then one workspace.
__host__ __device__ inline size_t rec_shared_doubles(int N, int D, int nw) {
    return rec_ring_doubles(N, D) + rec_lg_doubles(N) + rec_dir_doubles(N, nw);
}

__device__ __forceinline__ void rec_carve_dir(RecShared& s, double*& p, int N, int nw) {
    s.U = p; p += N + 1;
    s.R = p; p += REC_B;
    s.z = p; p += REC_B;
    s.part = p; p += (size_t)nw * REC_B;
    s.psum = p; p += (size_t)nw * REC_B;
    s.a = p; p += (size_t)REC_TRI;
}

// The ring, ln i, then the workspace; nw: the waves of the group that will
// run the recursion.
__device__ __forceinline__ RecShared rec_carve(double* base, int N, int D, int nw) {
    RecShared s;
    double* p = base;
    s.x1 = p; p += (size_t)D * N;
    s.xP = p; p += (size_t)D * N;
    s.Shi = p; p += N;
    s.Slo = p; p += N;
    s.lg = p; p += N + 1;
    rec_carve_dir(s, p, N, nw);
    return s;
}

// Copies beads 1 and P into the workspace and builds the prefix sums there
// (whole workgroup); optionally also writes the sums to global memory.
template <int D>
__device__ __forceinline__ Ring ring_stage(const RecShared& sh, int N, const double* x1, const double* xP, double k,
                                           bool pbc, double L, double* Shi_out, double* Slo_out) {
    for (int e = threadIdx.x; e < D * N; e += blockDim.x) {
        sh.x1[e] = x1[e];
        sh.xP[e] = xP[e];
    }
    for (int i = threadIdx.x; i <= N; i += blockDim.x) sh.lg[i] = i > 0 ? log((double)i) : 0.0;
    __syncthreads();
    const Ring r = make_ring(sh.x1, sh.xP, sh.Shi, sh.Slo, k, pbc, L);
    if (threadIdx.x < WAVE) ring_prefix<D>(r, N, threadIdx.x, sh.Shi, sh.Slo);
    __syncthreads();
    // (after the barrier: the sums of other lanes' chunks are read, which in
    // a device-memory workspace needs the workgroup's ordering)
    if (Shi_out) {
        for (int v = threadIdx.x; v < N; v += blockDim.x) {
            Shi_out[v] = sh.Shi[v];
            Slo_out[v] = sh.Slo[v];
        }
    }
    return r;
}

// The target side of w(n, q) held by one lane: forward, n-1 is the cycle's
// last particle v (xP[v], S_v) and ln n is added; backward, N-n is its first
// particle u (x1[u], S_u).
template <int D>
struct RecTarget {
    double x[D];
    double shi, slo, lg;
};

template <int DIR, int D>
__device__ __forceinline__ RecTarget<D> rec_target(const RecShared& sh, int N, int n) {
    RecTarget<D> t;
    const int idx = DIR == 0 ? n - 1 : N - n;
    const double* xs = DIR == 0 ? sh.xP : sh.x1;
#pragma unroll
    for (int a = 0; a < D; ++a) t.x[a] = xs[a * N + idx];
    t.shi = sh.Shi[idx];
    t.slo = sh.Slo[idx];
    t.lg = DIR == 0 ? sh.lg[n] : 0.0;
    return t;
}

// w(n,q) (see file header) for the lane's target and source q.
template <int DIR, int D>
__device__ __forceinline__ double rec_w(const RecShared& sh, const Ring& r, double bt, int N, const RecTarget<D>& t,
                                        int q) {
    const int idx = DIR == 0 ? q : N - 1 - q;
    const double* xs = DIR == 0 ? sh.x1 : sh.xP;
    double d[D];
#pragma unroll
    for (int a = 0; a < D; ++a) d[a] = ring_image(xs[a * N + idx] - t.x[a], r);
    // beta~ ecyc = (beta~ k/2) (S_v - S_u + |d|^2)
    const double sd = DIR == 0 ? ring_sdiff(t.shi, t.slo, sh.Shi[idx], sh.Slo[idx])
                               : ring_sdiff(sh.Shi[idx], sh.Slo[idx], t.shi, t.slo);
    return fma(bt * r.halfk, sd + norm2_fma<D>(d), DIR == 0 ? t.lg : sh.lg[N - q]);
}

// Entry (t, j), j < t, of the strict lower triangle of a block taken by lane
// `lane` of virtual row vr (0 <= vr < REC_B/2): rows vr+1 and REC_B-1-vr
// together hold REC_B entries, so REC_B/2 virtual rows cover the triangle
// once (the last one half full; t = -1 for its idle lanes).
__device__ __forceinline__ void rec_tri(int vr, int lane, int& t, int& j) {
    if (lane <= vr) {
        t = vr + 1;
        j = lane;
    } else {
        t = REC_B - 1 - vr;
        j = lane - vr - 1;
        if (vr == REC_B / 2 - 1) t = -1;
    }
}

// The sequential in-block pass of one wave, lane t holding target t's value
// v_t: for j = 0 .. REC_B-2 in turn, every lane takes
//     v_t = min(v_t, a[j][t] + v_j)   (KIND 0, the min-plus pass) or
//     v_t = v_t + a[j][t] v_j         (KIND 1, the forward substitution),
// v_j being what lane j holds at step j (final by then).  Target t's result is
// its value at step t: the broadcast v_j is stored to out[j] as it is taken
// (out[REC_B-1] from the last lane after the last step).  After its own step a
// lane's value is no longer used, so the pass reads column t of the packed
// triangle straight on past its end (into the next columns: finite entries
// or +inf, in bounds for every t) instead of selecting +inf / 0 there.
// Straight-line code, with a[j][t] read REC_PF steps ahead, so each step
// costs one cross-lane read and one or two operations on the dependent chain.
// The results are in out[] for the whole wave on return.
template <int KIND>
__device__ __forceinline__ void rec_inblock_seq(const double* a, int t, double v, double* out) {
    const int base = rec_idx(0, t);   // rec_idx(REC_B - 2, REC_B - 1) = REC_TRI - 1 bounds every read
    double buf[REC_PF];
#pragma unroll
    for (int u = 0; u < REC_PF; ++u) buf[u] = a[base + u];
#pragma unroll
    for (int j = 0; j < REC_B - 1; ++j) {
        const double aj = buf[j % REC_PF];
        if (j + REC_PF < REC_B - 1) buf[j % REC_PF] = a[base + j + REC_PF];
        const double vj = wave_bcast_uniform(v, j);
        out[j] = vj;
        v = KIND == 0 ? fmin(v, aj + vj) : fma(aj, vj, v);
    }
    if (t == REC_B - 1) out[REC_B - 1] = v;
    __builtin_amdgcn_wave_barrier();
}

// Blocks of the recursions for N particles; the backward recursion has N-1
// targets and may leave the last block empty.
__host__ __device__ inline int rec_blocks(int N) { return (N + REC_B - 1) / REC_B; }

// One recursion (DIR 0: forward V, DIR 1: backward Vb) for one ring on the
// staged ring (ring_stage), run by a group of gnthr threads (a multiple of 64)
// of the workgroup; gtid is the thread's index in its group.  Every thread of
// the workgroup calls this (or the other direction, in a second group) and
// passes the same workgroup barriers: one after set-up, four per block for
// rec_blocks(N) blocks, one at the end.  Lane t of every wave serves target
// n0+t of the current block; the group's waves split the sources of earlier
// blocks and keep UNROLL evaluations in flight.  Writes V[0..N] and Vb[0]
// (DIR 0) or Vb[1..N] (DIR 1); sets *status to 1 when a potential is not
// finite (the reference throws overflow_error,
// quadratic_bosonic_exchange.cpp:92-97, 119-124).
template <int DIR, int D, int UNROLL = REC_UNROLL>
__device__ void exchange_recursion(int N, double bt, const Ring& r, double* __restrict__ V, double* __restrict__ Vb,
                                   int* __restrict__ status, RecShared sh, int gtid, int gnthr) {
    const int tid = gtid, nthr = gnthr;
    // the wave index in a scalar register: loops over it are uniform
    const int lane = tid % WAVE, wave = __builtin_amdgcn_readfirstlane(tid / WAVE);
    // the partials are sized for the group's waves (rec_carve)
    const int nwave = nthr / WAVE;
    const int ntarget = DIR == 0 ? N : N - 1;
    for (int i = tid; i <= N; i += nthr) sh.U[i] = 0.0;   // (ln i is staged with the ring)
    __syncthreads();
    const int nblk = rec_blocks(N);
    for (int blk = 0; blk < nblk; ++blk) {
        const int n0 = 1 + blk * REC_B;
        const int nb = min(REC_B, ntarget - n0 + 1);
        if (nb <= 0) {   // the backward recursion's empty last block: barriers only
            __syncthreads();
            __syncthreads();
            __syncthreads();
            __syncthreads();
            continue;
        }
        const int tn = n0 + min(lane, nb - 1);   // the lane's target (clamped)
        const RecTarget<D> tg = rec_target<DIR, D>(sh, N, tn);
        // (0) in-block weights w(n0+t, n0+j), j < t < nb, staged as a[j][t]
        // (the strict lower triangle, two rows per wave pass); +inf for
        // targets past the block's end
        for (int vr = wave; vr < REC_B / 2; vr += nthr / WAVE) {
            int t, j;
            rec_tri(vr, lane, t, j);
            if (t >= 0 && t < nb) {
                const RecTarget<D> tt = rec_target<DIR, D>(sh, N, n0 + t);
                sh.a[rec_idx(j, t)] = rec_w<DIR, D>(sh, r, bt, N, tt, n0 + j);
            } else if (t >= nb) {
                sh.a[rec_idx(j, t)] = INFINITY;
            }
        }
        // (1) one pass over this wave's share of earlier sources: the lane's
        // running minimum m of y_q = w(n,q) + U_q and s = sum_q exp(-(y_q - m)),
        // rescaled whenever a chunk lowers m
        const int chunk = (n0 + nwave - 1) / nwave;
        const int qlo = min(n0, wave * chunk), qhi = min(n0, qlo + chunk);
        const bool active = wave < nwave && lane < nb;
        if (wave < nwave) {
            double m = INFINITY, s = 0.0;
            if (active) {
                int q = qlo;
                for (; q + UNROLL <= qhi; q += UNROLL) {
                    double y[UNROLL];
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) y[u] = rec_w<DIR, D>(sh, r, bt, N, tg, q + u) + sh.U[q + u];
                    double cm = y[0];
#pragma unroll
                    for (int u = 1; u < UNROLL; ++u) cm = fmin(cm, y[u]);
                    const double mn = fmin(m, cm);
                    s *= exp_pair(mn - m);   // m = inf (first chunk): s = 0 stays 0
                    m = mn;
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) s += exp_pair(mn - y[u]);
                }
                for (; q < qhi; ++q) {
                    const double y = rec_w<DIR, D>(sh, r, bt, N, tg, q) + sh.U[q];
                    const double mn = fmin(m, y);
                    s = fma(s, exp_pair(mn - m), exp_pair(mn - y));
                    m = mn;
                }
            }
            sh.part[wave * REC_B + lane] = m;
            sh.psum[wave * REC_B + lane] = s;
        }
        __syncthreads();
        // (1b) in-block min-plus on wave 0 (a[j][t] = +inf for j >= t)
        if (wave == 0) {
            const int t = lane;
            double rr = INFINITY;
            for (int w = 0; w < nwave; ++w) rr = fmin(rr, sh.part[w * REC_B + t]);
            if (t >= nb) rr = INFINITY;
            rec_inblock_seq<0>(sh.a, t, rr, sh.R);
        }
        __syncthreads();
        // (2a) each wave's earlier-source sum relative to R_t,
        // s_w exp(-(m_w - R_t)), in place of s_w (added in wave order by (2c))
        if (wave < nwave) {
            const int t = lane;
            const double sw = sh.psum[wave * REC_B + t];
            double c = 0.0;
            if (t < nb && sw > 0.0) c = sw * exp_pair(sh.R[t] - sh.part[wave * REC_B + t]);
            sh.psum[wave * REC_B + t] = c;
        }
        // (2b) in-block coefficients a[j][t] = exp(-(w(n0+t, n0+j) + R_j - R_t)),
        // j < t < nb; 0 elsewhere
        for (int vr = wave; vr < REC_B / 2; vr += nthr / WAVE) {
            int t, j;
            rec_tri(vr, lane, t, j);
            if (t >= 0 && t < nb) sh.a[rec_idx(j, t)] = exp_pair(-(sh.a[rec_idx(j, t)] + sh.R[j] - sh.R[t]));
            else if (t >= nb) sh.a[rec_idx(j, t)] = 0.0;
        }
        __syncthreads();
        // (2c) in-block forward substitution on wave 0 (a[j][t] = 0 for
        // j >= t), then U = R - ln z
        if (wave == 0) {
            const int t = lane;
            double z0 = 0.0;
            for (int w = 0; w < nwave; ++w) z0 += sh.psum[w * REC_B + t];
            rec_inblock_seq<1>(sh.a, t, z0, sh.z);
            const double z = sh.z[t];
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

// Block 0 runs the forward recursion (and writes the ring's prefix sums for
// the exterior forces), block 1 the backward one, each on the workspace at
// base (rec_shared_doubles(N, D, waves) doubles).  x1, xP: SoA [a][i].
template <int D>
__device__ __forceinline__ void exchange_recursions_on(double* base, int N, double bt, double k, int pbc, double L,
                                                       const double* __restrict__ x1, const double* __restrict__ xP,
                                                       double* __restrict__ V, double* __restrict__ Vb,
                                                       double* __restrict__ Shi, double* __restrict__ Slo,
                                                       int* __restrict__ status) {
    const RecShared sh = rec_carve(base, N, D, blockDim.x / WAVE);
    const bool fwd = blockIdx.x == 0;
    const Ring r = ring_stage<D>(sh, N, x1, xP, k, pbc != 0, L, fwd ? Shi : nullptr, fwd ? Slo : nullptr);
    if (fwd) exchange_recursion<0, D>(N, bt, r, V, Vb, status, sh, threadIdx.x, blockDim.x);
    else exchange_recursion<1, D>(N, bt, r, V, Vb, status, sh, threadIdx.x, blockDim.x);
}

// Both recursions with the workspace in LDS (rec_shared_doubles doubles of
// dynamic shared memory).
template <int D>
static __global__ void __launch_bounds__(1024) k_exchange_recursions(int N, double bt, double k, int pbc, double L,
                                                                     const double* __restrict__ x1,
                                                                     const double* __restrict__ xP,
                                                                     double* __restrict__ V, double* __restrict__ Vb,
                                                                     double* __restrict__ Shi, double* __restrict__ Slo,
                                                                     int* __restrict__ status) {
    extern __shared__ double smem[];
    exchange_recursions_on<D>(smem, N, bt, k, pbc, L, x1, xP, V, Vb, Shi, Slo, status);
}

// The same for rings whose workspace exceeds a workgroup's LDS: block b's
// workspace at ws + b rec_shared_doubles(N, D, waves), in device memory.  The
// workspace is only touched between workgroup barriers or by the lane that
// wrote it (the sequential passes read back their own stores), so the
// workgroup's ordering of device memory is all the recursion relies on.
template <int D>
static __global__ void __launch_bounds__(1024) k_exchange_recursions_gws(int N, double bt, double k, int pbc,
                                                                         double L, const double* __restrict__ x1,
                                                                         const double* __restrict__ xP,
                                                                         double* __restrict__ V,
                                                                         double* __restrict__ Vb,
                                                                         double* __restrict__ Shi,
                                                                         double* __restrict__ Slo,
                                                                         int* __restrict__ status,
                                                                         double* __restrict__ ws) {
    double* base = ws + (size_t)blockIdx.x * rec_shared_doubles(N, D, blockDim.x / WAVE);
    exchange_recursions_on<D>(base, N, bt, k, pbc, L, x1, xP, V, Vb, Shi, Slo, status);
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
// Probabilities are recomputed from V, Vb and the ring (conn_prob's
// expression, :142-157); the displacement of a term also gives its cycle
// energy (ecyc(l,j) for bead 1, ecyc(j,l) for bead P).  x2 / xPm1 hold the
// interior neighbours; with grad_only the interior link is left out and the
// result is -dV_B/dx (the adjoint of V_B alone, used by the autograd harness).
template <int D, int JUNROLL = 4>
__device__ __forceinline__ void exterior_force(int which, int l, int N, int lane, double bt, double k, const Ring& r,
                                               const double* __restrict__ V, const double* __restrict__ Vb,
                                               const double* __restrict__ x2, const double* __restrict__ xPm1,
                                               bool grad_only, double (&F)[D]) {
    const double* self = which == 0 ? r.x1 : r.xP;
    const double* other = which == 0 ? r.xP : r.x1;
    double xs[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xs[a] = self[a * N + l];
    const double shl = r.Shi[l], sll = r.Slo[l];
    double S[D];
#pragma unroll
    for (int a = 0; a < D; ++a) S[a] = 0.0;
    const int jlo = which == 0 ? max(0, l - 1) : 0;
    const int jhi = which == 0 ? N - 1 : min(l + 1, N - 1);
    const double VN = V[N];
#pragma unroll JUNROLL
    for (int j = jlo + lane; j <= jhi; j += WAVE) {
        const int row = which == 0 ? j : l, col = which == 0 ? l : j;
        double d[D];
#pragma unroll
        for (int a = 0; a < D; ++a) d[a] = ring_image(other[a * N + j] - xs[a], r);
        const bool link = col == row + 1 && row < N - 1;
        double pr = 0.0;
        if (link || col <= row) {
            double arg;
            if (link) {
                arg = V[row + 1] + Vb[row + 1] - VN;
            } else {
                // ecyc(col, row): u = col, v = row
                const double e = which == 0 ? ring_ecyc<D>(r, d, shl, sll, r.Shi[j], r.Slo[j])
                                            : ring_ecyc<D>(r, d, r.Shi[j], r.Slo[j], shl, sll);
                arg = V[col] + e + Vb[row + 1] - VN;
            }
            const double ex = exp(-bt * arg);
            pr = link ? 1.0 - ex : 1.0 / (row + 1) * ex;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) S[a] += pr * d[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) S[a] = wave_sum(S[a]);
    const double* link = which == 0 ? x2 : xPm1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double s = S[a];
        if (!grad_only) s += ring_image(link[a * N + l] - xs[a], r);
        F[a] = s * k;
    }
}

// Adds the exterior spring forces into f for beads 1 and P:
// f = F_exchange + f_phys (simulation.cpp:363-373: spring + physical).
// Shi/Slo: the ring's prefix sums from k_exchange_recursions.
template <int D>
static __global__ void __launch_bounds__(256) k_exchange_forces(int N, int P, double bt, double k, int pbc, double L,
                                                         const double* __restrict__ Shi,
                                                         const double* __restrict__ Slo,
                                                         const double* __restrict__ V,
                                                         const double* __restrict__ Vb, const double* __restrict__ x,
                                                         double* __restrict__ f, double* __restrict__ fex) {
    const int task = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (task >= 2 * N) return;
    const int which = task / N, l = task % N;
    const int lane = threadIdx.x % WAVE;
    const size_t bs = (size_t)D * N;
    const double* x2 = x + bs * (1 % P);
    const double* xPm1 = x + bs * ((P - 2 + P) % P);
    const Ring r = make_ring(x, x + bs * (P - 1), Shi, Slo, k, pbc != 0, L);
    double F[D];
    exterior_force<D>(which, l, N, lane, bt, k, r, V, Vb, x2, xPm1, false, F);
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
                                                          const double* __restrict__ Shi,
                                                          const double* __restrict__ Slo,
                                                          const double* __restrict__ V,
                                                          const double* __restrict__ Vb, const double* __restrict__ x1,
                                                          const double* __restrict__ xP, double* __restrict__ out) {
    const int task = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (task >= 2 * N) return;
    const int which = task / N, l = task % N;
    const int lane = threadIdx.x % WAVE;
    const Ring r = make_ring(x1, xP, Shi, Slo, k, pbc != 0, L);
    double F[D];
    exterior_force<D>(which, l, N, lane, bt, k, r, V, Vb, x1, xP, true, F);
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) out[((size_t)which * D + a) * N + l] = F[a];
    }
}

}  // namespace pimdb
