// exchange_batch.hpp -- the exchange of one ring on one workgroup, as the
// batched form runs it: the forward recursion, then a single pass that solves
// the backward recursion in linear form and accumulates the exterior spring
// forces of beads 1 and P from the same terms.
//
// Reference: src/bosonic_exchange/quadratic_bosonic_exchange.cpp of
// higj/pimd-b: cycle energies :34-61, forward V :73-99, backward Vb :101-128,
// connection probabilities :142-157, exterior forces :159-215.
//
// Exponents in base 2.  Positions are staged in LDS in units of the box with
// periodic boundaries (the minimum image of a displacement t is then
// t - rint(t)) and as they are without.  With c2 = beta~ k/2 log2(e) (times
// L^2 with periodic boundaries), S_v = sum_{j<v} |x1[j+1] - xP[j]|^2 the link
// prefix sums of the staged positions and U_n = beta~ V[n] log2(e), the
// forward recursion (:73-99) is
//     U_n = -log2 sum_{q<n} 2^-y(n,q),
//     y(n,q) = c2 (S_{n-1} - S_q + |xP[n-1] - x1[q]|^2) + log2 n + U_q,
// since beta~ E(q..n-1) = (beta~ k/2)(S_{n-1} - S_q + |x1[q] - xP[n-1]|^2)
// (the closed form of the reference's recurrence, :45-60).
//
// Rounding.  S is held split as S = H + L (exchange.hpp, Ring): H a multiple
// of a power of two near 2^-40 S_N, so H_v - H_q is exact, and |L| below half
// that quantum, so c2 L is of order 1e-8 in the exponents' units.  The
// difference H_v - H_q stays in the per-pair arithmetic, which keeps every
// short cycle's energy accurate to its own size; c2 L goes into per-particle
// terms together with U and log2 n:
//     y(n,q) = c2 (H_{n-1} - H_q + |xP[n-1] - x1[q]|^2) + [log2 n + c2 L_{n-1}] + [U_q - c2 L_q],
// rounded at the scale of U, as the reference's own sums V[u] + E + Vb - V[N]
// are.  Folding c2 H as well would round every short cycle at the scale of
// c2 S_N, far above the ring's free energy when the exchange links are long
// (1e4 against 1e3 for the benchmark's bead-replicated start), and the error
// would grow along the recursion.
//
// The forward recursion runs in blocks of 64 targets as exchange_recursion
// does (an online log-sum-exp over the earlier blocks' sources on every wave,
// then the block's min-plus pass and forward substitution on one wave).
//
// Backward recursion in linear form.  With G_l = exp(-beta~ (V[l] + Vb[l] -
// V[N])) for l = 0..N (G_N = 1, and G_0 = 1 since Vb[0] = V[N]), the backward
// recursion (:101-128) multiplied by exp(-beta~ (V[l] - V[N])) reads
//     G_a = sum_{b >= a} c(a,b) G_{b+1},
//     c(a,b) = exp(-beta~ (E(a..b) + V[a] - V[b+1])) / (b+1)
//            = 2^( [U_{b+1} - log2(b+1) - c2 L_b] + [c2 L_a - U_a] - c2 (H_b - H_a + |xP[b] - x1[a]|^2) ),
// and each term is a connection probability (:142-157):
//     P[b][a] = c(a,b) G_{b+1},   P[a-1][a] = 1 - G_a.
// For fixed b the c(a,b) sum to 1 over a (the forward recursion), so every
// c <= 1 and 0 <= G <= 1: nothing overflows, and a G that underflows is a
// probability the reference's exp underflows as well.  The exterior forces
// (:159-215) add P[b][a] d, d = xP[b] - x1[a] (min image), to F_1[a] and
// subtract it from F_P[b].  So the pass that solves for G accumulates the
// forces with the same terms: two O(N^2) passes with an exponential per pair
// (forward; backward with the forces) instead of three.
//
// The backward pass takes target blocks A of 64 particles from the last to the
// first.  Round A:
//   * chunks, one per wave: the off-diagonal tiles (A, B > A), whose G_{b+1}
//     are known from earlier rounds, cut into chunks of lane offsets (lane t
//     holds particle a = 64A + t and meets b = 64B + (t + s) mod 64 at offset
//     s; h_a and its F_1 partial stay in registers, the F_P partial of its
//     column moves one lane along the wave after every offset, DPP), and the
//     forces of the diagonal tile of block A+1 (its coefficients are stored);
//     each chunk leaves its partials in its own LDS slot, added per particle
//     in chunk order afterwards (deterministic);
//   * the coefficients c(a,b), a <= b, of the diagonal tile of block A, packed
//     by row;
//   * one wave solves the block: G_a = h_a + c(a, last) G_next + sum over the
//     block's b >= a of c(a,b) G_{b+1}, an upper-triangular system taken from
//     the last target down, one FMA and one cross-lane read per target on the
//     dependent chain.
// A last round adds the forces of block 0's diagonal tile; then the links
// b = a - 1 are added per particle.
#pragma once

#include "exchange.hpp"

namespace pimdb {

// The diagonal tile's upper triangle (with the diagonal), packed by row:
// entry (t, j), t <= j < 64, at xb_row(t) + (j - t).
constexpr int XB_DIAG = REC_B * (REC_B + 1) / 2;
__host__ __device__ __forceinline__ int xb_row(int t) { return t * REC_B - t * (t - 1) / 2; }
// Forward source record of particle q: x1[q] (D doubles), H_q, U_q - c2 L_q
// (written when U_q is known); read by whole waves at one address (a
// broadcast), so its stride is only kept even.
__host__ __device__ constexpr int xb_frec(int D) { return (D + 3) & ~1; }
// Backward source record of particle b: xP[b] (D doubles), H_b,
// U_{b+1} - log2(b+1) - c2 L_b, G_{b+1}; an odd stride, so that lanes reading
// consecutive records spread over the LDS banks.
__host__ __device__ constexpr int xb_rec(int D) { return (D + 3) | 1; }
// One chunk's partials: h (64), F_1 (D x 64), F_P (D x 64).
__host__ __device__ constexpr int xb_slot(int D) { return (1 + 2 * D) * REC_B; }
// Particles per thread when the backward records are built (N <= 4 threads).
constexpr int XB_MAXPER = 4;

// LDS layout (offsets in doubles).  x1, H and L live through both passes (L
// becomes the backward target term c2 L_a - U_a); the backward workspace
// overlays U and everything after it.
struct XbLayout {
    int x1, Shi, Slo;                 // bead 1 [D][N], the split sums H [N], L [N]
    int U2, xP, lg2, frec;            // U [N+1], bead P [D][N], log2 n [N+1], source records [N]
    int R, z, part, psum, tri;        // forward: shifts, substitution results, partials, in-block matrix
    int rec, acc, hsum, dbuf, slot;   // backward: records [N], sums (F_1 [2][D][64], F_P [D][N]), h [64] + [64], diagonal tile, chunk slots
    int total;
};

__host__ __device__ inline XbLayout xb_layout(int N, int D, int nw, int nslot) {
    XbLayout l;
    int p = 0;
    l.x1 = p; p += D * N;
    l.Shi = p; p += N;
    l.Slo = p; p += N;
    const int bwd = p;
    l.U2 = p; p += N + 1;
    l.xP = p; p += D * N;
    l.lg2 = p; p += N + 1;
    p += p & 1;
    l.frec = p; p += N * xb_frec(D);
    l.R = p; p += REC_B;
    l.z = p; p += REC_B;
    l.part = p; p += nw * REC_B;
    l.psum = p; p += nw * REC_B;
    l.tri = p; p += REC_TRI;
    const int fwd_end = p;
    p = bwd;
    l.rec = p; p += N * xb_rec(D);
    l.acc = p; p += 2 * D * REC_B + D * N;
    l.hsum = p; p += 2 * REC_B;   // h, then the solve's broadcast values
    l.dbuf = p; p += XB_DIAG;
    l.slot = p; p += nslot * xb_slot(D);
    l.total = fwd_end > p ? fwd_end : p;
    return l;
}

// Minimum image of a staged displacement (box units) or the displacement itself.
template <bool PBC>
__device__ __forceinline__ double xb_image(double t) {
    if constexpr (PBC) return t - rint(t);
    else return t;
}

// Squared (min-imaged) distance between staged particles.
template <int D, bool PBC>
__device__ __forceinline__ double xb_dist2(const double* xa, int ia, const double* xb, int ib, int N) {
    double d[D];
#pragma unroll
    for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(xa[ax * N + ia] - xb[ax * N + ib]);
    return norm2_fma<D>(d);
}

// The split link prefix sums of the staged positions by one wave: chunks of
// ceil(N/64) links per lane summed as double-doubles, an exclusive scan of
// the chunk sums across lanes, then the running sums (the chunk's links are
// evaluated again), stored as S_v = H_v + L_v with H_v a multiple of
// q = 2^(e - 40), 2^e <= S_{N-1} < 2^(e+1) (as ring_prefix): H_v - H_u is exact
// and |L_v| <= q/2.  H_v also goes into the forward source records.
template <int D, bool PBC>
__device__ __forceinline__ void xb_prefix(const double* x1, const double* xP, int N, int lane, double* Shi,
                                          double* Slo, double* frec) {
    const int c = (N + WAVE - 1) / WAVE;
    const int v0 = lane * c;
    DD tot{0.0, 0.0};
    for (int t = 0; t < c; ++t) {
        const int v = v0 + t;
        if (v < N - 1) tot = dd_add(tot, DD{xb_dist2<D, PBC>(x1, v + 1, xP, v, N), 0.0});
    }
    DD inc = tot;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const DD o{__shfl_up(inc.hi, off, WAVE), __shfl_up(inc.lo, off, WAVE)};
        if (lane >= off) inc = dd_add(o, inc);
    }
    DD run{__shfl_up(inc.hi, 1, WAVE), __shfl_up(inc.lo, 1, WAVE)};
    if (lane == 0) run = DD{0.0, 0.0};
    const double total = __shfl(inc.hi, WAVE - 1, WAVE);   // lane 63's inclusive sum: S_{N-1}
    const double qq = total > 0.0 ? ldexp(1.0, ilogb(total) - 40) : 1.0;
    const double invq = 1.0 / qq;
    for (int t = 0; t < c; ++t) {
        const int v = v0 + t;
        if (v < N) {
            const double h = rint(run.hi * invq) * qq;
            Shi[v] = h;
            Slo[v] = (run.hi - h) + run.lo;
            frec[v * xb_frec(D) + D] = h;
            if (v < N - 1) run = dd_add(run, DD{xb_dist2<D, PBC>(x1, v + 1, xP, v, N), 0.0});
        }
    }
}

// The forward recursion on the whole workgroup (all threads call it; four
// barriers per block of 64 targets).  Writes U[1..N] and the records' U terms;
// sets *status when a potential is not finite (the reference throws
// overflow_error, quadratic_bosonic_exchange.cpp:92-97).
//
// A block's sources q < n0 are split in two.  Those of the blocks before the
// previous one were taken early: while wave 0 ran the previous block's
// min-plus pass and substitution, waves 1 .. nw-1 summed them for this
// block's targets (half beside each pass), and each keeps its running
// (m, s) in registers.  At the start of the block every wave takes its
// share of the previous block's own targets, whose U was not known then,
// continuing from those registers; the per-wave partials then enter the
// min-plus shift and the sum as before.  The barriers per block stay four
// and the layout is unchanged.
template <int D, bool PBC, int UNROLL>
__device__ __forceinline__ void xb_forward(int N, double c2, const XbLayout& lay, double* lds, int* status) {
    constexpr int FR = xb_frec(D);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid % WAVE, wave = __builtin_amdgcn_readfirstlane(tid / WAVE), nw = nthr / WAVE;
    const double* x1 = lds + lay.x1;
    const double* xP = lds + lay.xP;
    const double* Shi = lds + lay.Shi;
    const double* Slo = lds + lay.Slo;
    const double* lg2 = lds + lay.lg2;
    double* U = lds + lay.U2;
    double* frec = lds + lay.frec;
    double* R = lds + lay.R;
    double* zz = lds + lay.z;
    double* part = lds + lay.part;
    double* psum = lds + lay.psum;
    double* a = lds + lay.tri;
    const int nblk = rec_blocks(N);
    const bool early = nw > 1;     // waves besides the one running the passes
    // The lane's target of the block starting at n0 (clamped past the block's
    // end): xP[n-1], H_{n-1}, log2 n + c2 L_{n-1}.
    struct Tgt {
        double x[D], h, T;
    };
    auto target = [&](int n0, int nb) {
        Tgt g;
        const int tn = n0 + min(lane, nb - 1);
#pragma unroll
        for (int ax = 0; ax < D; ++ax) g.x[ax] = xP[ax * N + tn - 1];
        g.h = Shi[tn - 1];
        g.T = fma(c2, Slo[tn - 1], lg2[tn]);
        return g;
    };
    // sources [qlo, qhi) into the lane's running minimum m of y and
    // s = sum 2^(m - y), in groups of UNROLL (the last group padded with
    // y = +inf, whose 2^(m - y) is taken as 2^-2000 = 0), rescaled when a
    // group lowers m
    auto accumulate = [&](const Tgt& g, int qlo, int qhi, double& m, double& s) {
        auto ysrc = [&](int q) {
            const double* fr = frec + q * FR;
            double d[D];
#pragma unroll
            for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(g.x[ax] - fr[ax]);
            return fma(c2, (g.h - fr[D]) + norm2_fma<D>(d), g.T + fr[D + 1]);
        };
        int q0 = qlo;
        for (; q0 + UNROLL <= qhi; q0 += UNROLL) {
            // the group's records first (all loads in flight together)
            double rv[UNROLL][D + 2];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
#pragma unroll
                for (int f = 0; f < D + 2; ++f) rv[u][f] = frec[(q0 + u) * FR + f];
            double y[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                double d[D];
#pragma unroll
                for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(g.x[ax] - rv[u][ax]);
                y[u] = fma(c2, (g.h - rv[u][D]) + norm2_fma<D>(d), g.T + rv[u][D + 1]);
            }
            double cm = y[0];
#pragma unroll
            for (int u = 1; u < UNROLL; ++u) cm = fmin(cm, y[u]);
            const double mn = fmin(m, cm);
            s *= exp2_pair(fmax(mn - m, -2000.0));   // m = +inf on the first group: s = 0 stays 0
            m = mn;
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) s += exp2_pair(mn - y[u]);
        }
        if (q0 < qhi) {   // the last, partial group
            double y[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) y[u] = q0 + u < qhi ? ysrc(q0 + u) : INFINITY;
            double cm = y[0];
#pragma unroll
            for (int u = 1; u < UNROLL; ++u) cm = fmin(cm, y[u]);
            const double mn = fmin(m, cm);
            s *= exp2_pair(fmax(mn - m, -2000.0));
            m = mn;
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) s += exp2_pair(fmax(mn - y[u], -2000.0));
        }
    };
    double em = INFINITY, es = 0.0;   // waves >= 1: this wave's early partials for the next block's targets
    for (int blk = 0; blk < nblk; ++blk) {
        const int n0 = 1 + blk * REC_B;
        const int nb = min(REC_B, N - n0 + 1);
        const Tgt tg = target(n0, nb);
        // sources of this block taken early (blocks before the previous one)
        const bool had_early = early && blk >= 1;
        const int lo1 = had_early ? n0 - REC_B : 0;
        // (0) in-block weights y(n0+t, n0+j) - U_{n0+j}, j < t < nb, at
        // a[rec_idx(j, t)]; +inf for targets past the block's end
        for (int vr = wave; vr < REC_B / 2; vr += nw) {
            int t, j;
            rec_tri(vr, lane, t, j);
            if (t >= 0) {
                double w = INFINITY;
                if (t < nb) {
                    const int n = n0 + t, q = n0 + j;
                    const double T = fma(c2, Slo[n - 1], lg2[n]);
                    w = fma(c2, (Shi[n - 1] - Shi[q]) + xb_dist2<D, PBC>(xP, n - 1, x1, q, N), fma(-c2, Slo[q], T));
                }
                a[rec_idx(j, t)] = w;
            }
        }
        // (1) this wave's share of the sources [lo1, n0) (a balanced
        // contiguous range)
        {
            const int span = n0 - lo1;
            const int qlo = lo1 + (int)(((long)span * wave) / nw), qhi = lo1 + (int)(((long)span * (wave + 1)) / nw);
            double m = INFINITY, s = 0.0;
            if (had_early && wave >= 1) {   // continue from the early sources
                m = em;
                s = es;
            }
            if (lane < nb) accumulate(tg, qlo, qhi, m, s);
            part[wave * REC_B + lane] = m;
            psum[wave * REC_B + lane] = s;
        }
        __syncthreads();
        // the next block's early sources [0, n0) on waves 1 .. nw-1, each
        // wave's range in two halves (beside the min-plus pass, then beside
        // the substitution)
        const bool give = early && blk + 1 < nblk;
        const int n1 = n0 + REC_B;
        const int nb1 = give ? min(REC_B, N - n1 + 1) : 0;
        em = INFINITY;
        es = 0.0;
        int elo = 0, emid = 0, ehi = 0;
        if (give && wave >= 1) {
            elo = (int)(((long)n0 * (wave - 1)) / (nw - 1));
            ehi = (int)(((long)n0 * wave) / (nw - 1));
            emid = elo + (ehi - elo) / 2;
        }
        // (1b) in-block min-plus on wave 0: R_t = min(min_w m_w, min_j a[j][t] + R_j)
        if (wave == 0) {
            double rr = INFINITY;
            for (int w = 0; w < nw; ++w) rr = fmin(rr, part[w * REC_B + lane]);
            if (lane >= nb) rr = INFINITY;
            rec_inblock_seq<0>(a, lane, rr, R);
        } else if (give && lane < nb1) {
            accumulate(target(n1, nb1), elo, emid, em, es);
        }
        __syncthreads();
        // (2a) each wave's earlier-source sums relative to R_t
        {
            const int t = lane;
            const double sw = psum[wave * REC_B + t];
            double c = 0.0;
            if (t < nb && sw > 0.0) c = sw * exp2_pair(R[t] - part[wave * REC_B + t]);
            psum[wave * REC_B + t] = c;
        }
        // (2b) in-block coefficients 2^-(a[j][t] + R_j - R_t) (<= 1), 0 past the block's end
        for (int vr = wave; vr < REC_B / 2; vr += nw) {
            int t, j;
            rec_tri(vr, lane, t, j);
            if (t >= 0 && t < nb) a[rec_idx(j, t)] = exp2_pair(-(a[rec_idx(j, t)] + R[j] - R[t]));
            else if (t >= nb) a[rec_idx(j, t)] = 0.0;
        }
        __syncthreads();
        // (2c) forward substitution on wave 0, then U = R - log2 z
        if (wave == 0) {
            const int t = lane;
            double z0 = 0.0;
            for (int w = 0; w < nw; ++w) z0 += psum[w * REC_B + t];
            rec_inblock_seq<1>(a, t, z0, zz);
            const double z = zz[t];
            if (t < nb) {
                const int n = n0 + t;
                const double u = R[t] - log2(z);
                U[n] = u;
                if (n < N) frec[n * FR + D + 1] = fma(-c2, Slo[n], u);
                if (!isfinite(u)) atomicOr(status, 1);
            }
        } else if (give && lane < nb1) {
            accumulate(target(n1, nb1), emid, ehi, em, es);
        }
        __syncthreads();
    }
}

// One chunk of an off-diagonal tile (rows of block A, columns of block B > A),
// offsets [s0, s0+len): h_a = sum_b c(a,b) G_{b+1}, the F_1 partials of the
// rows and the F_P partials of the columns (without k; box units with periodic
// boundaries) into the chunk's slot.  Tb: the backward target terms
// c2 L_a - U_a.  One wave.
template <int D, bool PBC>
__device__ __forceinline__ void xb_offdiag_chunk(int N, double c2, const double* x1, const double* Shi,
                                                 const double* Tb, const double* rec, int A, int B, int s0, int len,
                                                 double* sl) {
    constexpr int REC = xb_rec(D);
    const int lane = threadIdx.x % WAVE;
    const int a = A * REC_B + lane;   // < N: A is not the last block
    double xa[D], F1[D], FPr[D];
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
        xa[ax] = x1[ax * N + a];
        F1[ax] = 0.0;
        FPr[ax] = 0.0;
    }
    const double ha = Shi[a], ta = Tb[a];
    double h = 0.0;
    auto step = [&](int s, bool masked) {
        int b = B * REC_B + ((lane + s) & (WAVE - 1));
        bool ok = true;
        if (masked) {
            ok = b < N;
            b = ok ? b : N - 1;
        }
        const double* rp = rec + b * REC;
        double d[D];
#pragma unroll
        for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(rp[ax] - xa[ax]);
        double term = exp2_pair(fma(-c2, (rp[D] - ha) + norm2_fma<D>(d), rp[D + 1] + ta)) * rp[D + 2];
        if (masked) term = ok ? term : 0.0;
        h += term;
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            F1[ax] = fma(term, d[ax], F1[ax]);
            FPr[ax] = lane_rotate(fma(-term, d[ax], FPr[ax]));
        }
    };
    if (B * REC_B + REC_B <= N) {
        for (int s = s0; s < s0 + len; ++s) step(s, false);
    } else {
        for (int s = s0; s < s0 + len; ++s) step(s, true);
    }
    // lane t now holds the F_P partial of column (t + s0 + len) mod 64
    const int col = (lane + s0 + len) & (WAVE - 1);
    sl[lane] = h;
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
        sl[(1 + ax) * REC_B + lane] = F1[ax];
        sl[(1 + D + ax) * REC_B + col] = FPr[ax];
    }
}

// One chunk of the diagonal tile of block Ad (pairs a <= b within it): the
// probabilities c(a,b) G_{b+1} from the coefficients left in dbuf and the
// records' G; F_1 and F_P partials into the slot.
template <int D, bool PBC>
__device__ __forceinline__ void xb_diag_force_chunk(int N, const double* x1, const double* rec, const double* dbuf,
                                                    int Ad, int s0, int len, double* sl) {
    constexpr int REC = xb_rec(D);
    const int lane = threadIdx.x % WAVE;
    const int nbd = min(REC_B, N - Ad * REC_B);
    const int a = Ad * REC_B + min(lane, nbd - 1);
    double xa[D], F1[D], FPr[D];
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
        xa[ax] = x1[ax * N + a];
        F1[ax] = 0.0;
        FPr[ax] = 0.0;
    }
    const int rbase = xb_row(lane);
    for (int s = s0; s < s0 + len; ++s) {
        const int j = lane + s;
        const bool ok = j < nbd;   // (then lane < nbd and j < 64)
        const double* rp = rec + (Ad * REC_B + (ok ? j : 0)) * REC;
        double pr = dbuf[ok ? rbase + s : 0] * rp[D + 2];   // P[b][a] = c(a,b) G_{b+1}
        pr = ok ? pr : 0.0;
        double d[D];
#pragma unroll
        for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(rp[ax] - xa[ax]);
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            F1[ax] = fma(pr, d[ax], F1[ax]);
            FPr[ax] = lane_rotate(fma(-pr, d[ax], FPr[ax]));
        }
    }
    const int col = (lane + s0 + len) & (WAVE - 1);
    sl[lane] = 0.0;
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
        sl[(1 + ax) * REC_B + lane] = F1[ax];
        sl[(1 + D + ax) * REC_B + col] = FPr[ax];
    }
}

// Sum of one slot field (offset `off` in every slot) over chunks [lo, hi)
// of the sub-round starting at chunk c0, in chunk order.
__device__ __forceinline__ double xb_slot_sum(const double* slot, int sl, int off, int c0, int lo, int hi) {
    double s = 0.0;
#pragma unroll 4
    for (int c = lo; c < hi; ++c) s += slot[(c - c0) * sl + off];
    return s;
}

// The partials of chunks [c0, cend) of round A, added per particle in chunk
// order: h of block A into hsum (part 0), F_1 of block A (off-diagonal
// chunks, the first noff << lcpt) and of block A+1 (the diagonal-force
// chunks after them), F_P of the column tiles A+1 .. A+ncol (tile A+1+i's
// chunks, and for column tile A+1 also the diagonal-force chunks).  Parts
// [p0, p1) of 1 + 2D + D ncol (p1 < 0: all), by threads t0, t0 + ts, ...
// acc: F_1 of a row block at [(block & 1) D + axis][64] (a row block's sums
// are final one round after they start), then F_P at [axis][N].
template <int D>
__device__ __forceinline__ void xb_reduce(int N, int A, int noff, bool dforce, int lcpt, int c0, int cend,
                                          const double* slot, double* acc, double* hsum, int p0, int p1, int t0,
                                          int ts) {
    constexpr int SL = xb_slot(D);
    const int ncol = noff > 0 ? noff : (dforce ? 1 : 0);
    const int nparts = p1 < 0 ? 1 + 2 * D + D * ncol : p1;
    const int offend = min(cend, noff << lcpt);   // off-diagonal chunks end
    const int dlo = max(c0, noff << lcpt);          // diagonal-force chunks [dlo, cend)
    for (int e = p0 * REC_B + t0; e < nparts * REC_B; e += ts) {
        const int part = e / REC_B, l = e % REC_B;
        if (part == 0) {
            if (A >= 0) hsum[l] += xb_slot_sum(slot, SL, l, c0, c0, offend);
        } else if (part <= D) {
            const int ax = part - 1, row = A * REC_B + l;
            if (A >= 0 && row < N) acc[((A & 1) * D + ax) * REC_B + l] += xb_slot_sum(slot, SL, (1 + ax) * REC_B + l, c0, c0, offend);
        } else if (part <= 2 * D) {
            const int ax = part - 1 - D, row = (A + 1) * REC_B + l;
            if (dforce && row < N) acc[(((A + 1) & 1) * D + ax) * REC_B + l] += xb_slot_sum(slot, SL, (1 + ax) * REC_B + l, c0, dlo, cend);
        } else {
            const int q = part - 1 - 2 * D, i = q / D, ax = q % D, col = (A + 1 + i) * REC_B + l;
            if (col < N) {
                const int off = (1 + D + ax) * REC_B + l;
                double sum = 0.0;
                if (i < noff) sum = xb_slot_sum(slot, SL, off, c0, max(c0, i << lcpt), min(cend, (i + 1) << lcpt));
                if (i == 0) sum += xb_slot_sum(slot, SL, off, c0, dlo, cend);
                acc[2 * D * REC_B + ax * N + col] += sum;
            }
        }
    }
}

// Coefficients c(a,b), a <= b, of the diagonal tile of block A into dbuf
// (0 past the block's end), whole workgroup.  The 2080 pairs (t, t + s) are
// taken in 33 folded offsets s = 0..32: lane t takes (t, t + s) when
// t + s < 64 and otherwise (t + s - 64, t), the pair of offset 64 - s, so
// that every lane has a pair (at s = 32 the upper half would repeat the
// lower half's pairs and is idle).
template <int D, bool PBC>
__device__ __forceinline__ void xb_diag_coefs(int N, double c2, const double* x1, const double* Shi,
                                              const double* Tb, const double* rec, int A, double* dbuf) {
    constexpr int REC = xb_rec(D);
    const int lane = threadIdx.x % WAVE, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int nw = blockDim.x / WAVE;
    const int nbA = min(REC_B, N - A * REC_B);
    for (int s = wave; s <= REC_B / 2; s += nw) {
        const bool lower = lane + s < REC_B;
        if (s == REC_B / 2 && !lower) continue;
        const int r = lower ? lane : lane + s - REC_B;   // row (target) and column (source) in the block
        const int j = lower ? lane + s : lane;
        double cf = 0.0;
        if (j < nbA) {
            const int a = A * REC_B + r;
            const double* rp = rec + (A * REC_B + j) * REC;
            double d[D];
#pragma unroll
            for (int ax = 0; ax < D; ++ax) d[ax] = xb_image<PBC>(rp[ax] - x1[ax * N + a]);
            cf = exp2_pair(fma(-c2, (rp[D] - Shi[a]) + norm2_fma<D>(d), rp[D + 1] + Tb[a]));
        }
        dbuf[xb_row(r) + (j - r)] = cf;
    }
}

// The block's G on one wave: lane t holds h_t + c(t, last) G_next, then for
// u = 63 .. 1 every lane takes v_t += c(t, u-1) G_u with G_u broadcast from
// lane u (final by then: every later column has been added).  Lanes past the
// block's end hold 0 and add 0.  The broadcast G_u is stored to out[u] as it
// is taken (no per-lane conditions on the chain); lane t's value is not used
// after its own step, so its reads run on into the previous row for t >= u
// (in bounds), and the coefficients are read REC_PF steps ahead.  Then every
// lane t < nb writes G_t into record t - 1 (G_{b+1} of b = t - 1).
template <int D>
__device__ __forceinline__ void xb_solve(int N, int A, const double* hsum, const double* dbuf, double* rec,
                                         double* out) {
    constexpr int REC = xb_rec(D);
    const int t = threadIdx.x % WAVE;
    const int nbA = min(REC_B, N - A * REC_B);
    const int base = xb_row(t) - t - 1;   // c(t, u-1) at base + u
    // G of the next block's first target, or G_N = 1: record 64A + nb - 1
    const double gk = rec[(A * REC_B + nbA - 1) * REC + D + 2];
    double v = 0.0;
    if (t < nbA) v = fma(dbuf[base + nbA], gk, hsum[t]);
    double buf[REC_PF];
#pragma unroll
    for (int i = 0; i < REC_PF; ++i) buf[i] = dbuf[base + (REC_B - 1 - i)];
#pragma unroll
    for (int i = 0; i < REC_B - 1; ++i) {
        const int u = REC_B - 1 - i;
        const double cf = buf[i % REC_PF];
        if (i + REC_PF < REC_B - 1) buf[i % REC_PF] = dbuf[base + (u - REC_PF)];
        const double gu = wave_bcast_uniform(v, u);
        out[u] = gu;
        v = fma(cf, gu, v);
    }
    if (t == 0) out[0] = v;
    __builtin_amdgcn_wave_barrier();
    const int a = A * REC_B + t;
    if (t < nbA && a > 0) rec[(a - 1) * REC + D + 2] = out[t];
}

// The whole exchange of one ring on the workgroup: V (the reference's units)
// and the exterior spring forces added into beads 1 and P of f (f =
// F_exchange + f_phys, simulation.cpp:363-373), or written there (add false)
// for a caller that adds the physical forces afterwards.  x, f: bead-major SoA
// [P][D][N].  nslot: chunk slots of the layout (1 <= nslot <= waves).  Every
// thread calls it.
template <int D, bool PBC>
__device__ __forceinline__ void xb_exchange(int N, int P, double L, double k, double bt, const double* x, double* f,
                                            double* V, int* status, double* lds, int nslot, bool add) {
    constexpr int REC = xb_rec(D);
    constexpr int FR = xb_frec(D);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid % WAVE, wave = __builtin_amdgcn_readfirstlane(tid / WAVE), nw = nthr / WAVE;
    const XbLayout lay = xb_layout(N, D, nw, nslot);
    const size_t bs = (size_t)D * N;
    const double* x1g = x;
    const double* xPg = x + bs * (P - 1);
    const double sc = PBC ? 1.0 / L : 1.0;
    const double c2 = bt * 0.5 * k * 1.4426950408889634 * (PBC ? L * L : 1.0);
    double* x1 = lds + lay.x1;
    double* xP = lds + lay.xP;
    double* Shi = lds + lay.Shi;
    double* Slo = lds + lay.Slo;
    double* U = lds + lay.U2;
    double* lg2 = lds + lay.lg2;
    double* frec = lds + lay.frec;
    // ---- staging; the interior links' spring forces go into f now
    // (f = F_exchange + f_phys, simulation.cpp:363-373), so that the end
    // needs no other beads: f_1 += k (x_2 - x_1), f_P += k (x_{P-1} - x_P)
    // (min image in the reference's units, bosonic_exchange_base.cpp:72-78)
    const double* x2 = x + bs * (1 % P);
    const double* xPm1 = x + bs * ((P - 2 + P) % P);
    double* f1 = f;
    double* fP = f + bs * (P - 1);
    for (int e = tid; e < D * N; e += nthr) {
        const double s1 = x1g[e] * sc;
        x1[e] = s1;
        frec[(e % N) * FR + e / N] = s1;
        xP[e] = xPg[e] * sc;
    }
    for (int i = tid; i <= N; i += nthr) {
        lg2[i] = i > 0 ? log2((double)i) : 0.0;
        U[i] = 0.0;
    }
    __syncthreads();
    // the prefix sums on wave 0, the interior links on the others
    if (wave == 0) {
        xb_prefix<D, PBC>(x1, xP, N, lane, Shi, Slo, frec);
        if (lane == 0) frec[D + 1] = 0.0;   // U_0 - c2 L_0 = 0 (S_0 = 0)
    }
    if (wave != 0 || nw == 1) {
        const int t0 = nw == 1 ? tid : tid - WAVE, ts = nw == 1 ? nthr : nthr - WAVE;
        const double Lr = PBC ? L : 0.0, iLr = PBC ? 1.0 / L : 0.0;
        for (int e = t0; e < D * N; e += ts) {
            const double a1 = x1g[e], aP = xPg[e];
            const double d1 = x2[e] - a1, dP = xPm1[e] - aP;
            f1[e] = fma(fma(-Lr, rint(d1 * iLr), d1), k, add ? f1[e] : 0.0);
            fP[e] = fma(fma(-Lr, rint(dP * iLr), dP), k, add || P == 1 ? fP[e] : 0.0);   // (P = 1: the same bead)
        }
    }
    __syncthreads();
    // ---- forward recursion
    xb_forward<D, PBC, 4>(N, c2, lay, lds, status);
    // ---- backward pass with the exterior forces
    double* rec = lds + lay.rec;
    double* acc = lds + lay.acc;
    double* hsum = lds + lay.hsum;
    double* dbuf = lds + lay.dbuf;
    double* slot = lds + lay.slot;
    double* Tb = Slo;   // becomes c2 L_a - U_a
    {
        // V in the reference's units; then the records, which overlay U, xP,
        // lg2 and the forward workspace: read everything, then write
        const double vs = 6.931471805599453e-01 / bt;
        for (int i = tid; i <= N; i += nthr) V[i] = U[i] * vs;
        double rv[XB_MAXPER][D + 2];
#pragma unroll
        for (int r = 0; r < XB_MAXPER; ++r) {
            const int b = tid + r * nthr;
            if (b < N) {
#pragma unroll
                for (int ax = 0; ax < D; ++ax) rv[r][ax] = xP[ax * N + b];
                rv[r][D] = fma(-c2, Slo[b], U[b + 1] - lg2[b + 1]);
                rv[r][D + 1] = fma(c2, Slo[b], -U[b]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < XB_MAXPER; ++r) {
            const int b = tid + r * nthr;
            if (b < N) {
                double* rp = rec + b * REC;
#pragma unroll
                for (int ax = 0; ax < D; ++ax) rp[ax] = rv[r][ax];
                rp[D] = Shi[b];
                rp[D + 1] = rv[r][D];
                rp[D + 2] = b == N - 1 ? 1.0 : 0.0;   // G_N = 1; the others are filled as they are solved
                Tb[b] = rv[r][D + 1];
            }
        }
        for (int e = tid; e < 2 * D * REC_B + D * N; e += nthr) acc[e] = 0.0;
        if (tid < REC_B) hsum[tid] = 0.0;
        __syncthreads();
    }
    const int T = (N + REC_B - 1) / REC_B;
    double* out = hsum + REC_B;   // the solve's broadcast values
    const double kL = k * (PBC ? L : 1.0);
    // F_1 of row block B is final: with the links b = a - 1 (P[l-1][l] =
    // 1 - G_l, record l-1) into f, f_1 += k L (S + links) (the interior links
    // went in at the start), and its sums cleared for block B - 2
    auto flush_rows = [&](int B) {
        for (int e = tid; e < REC_B; e += nthr) {
            const int l = B * REC_B + e;
            double* s1 = acc + (B & 1) * D * REC_B + e;
            if (l < N) {
                double x1l[D], v[D];
#pragma unroll
                for (int ax = 0; ax < D; ++ax) {
                    x1l[ax] = x1[ax * N + l];
                    v[ax] = s1[ax * REC_B];
                }
                if (l >= 1) {
                    const double* rp = rec + (l - 1) * REC;
                    const double p1 = 1.0 - rp[D + 2];
                    if (isnan(p1)) atomicOr(status, 1);
#pragma unroll
                    for (int ax = 0; ax < D; ++ax) v[ax] = fma(p1, xb_image<PBC>(rp[ax] - x1l[ax]), v[ax]);
                }
#pragma unroll
                for (int ax = 0; ax < D; ++ax) {
                    const size_t i = (size_t)ax * N + l;
                    f1[i] = fma(v[ax], kL, f1[i]);
                }
            }
#pragma unroll
            for (int ax = 0; ax < D; ++ax) s1[ax * REC_B] = 0.0;
        }
    };
    for (int A = T - 1; A >= -1; --A) {
        // chunks: the off-diagonal tiles (A, B > A) and the diagonal tile of
        // block A+1 (solved in the previous round), 2^lcpt chunks per tile
        const int noff = A >= 0 ? T - 1 - A : 0;
        const bool dforce = A + 1 < T;
        const int ntiles = noff + (dforce ? 1 : 0);
        int lcpt = 0;   // at most 16 chunks per tile, one chunk per wave
        while (lcpt < 4 && (2 << lcpt) * ntiles <= nw) ++lcpt;
        const int len = REC_B >> lcpt, nchunk = ntiles << lcpt;
        int c0 = 0;
        for (; c0 < nchunk; c0 += nslot) {
            const int cend = min(nchunk, c0 + nslot);
            const int c = c0 + wave;
            if (c < cend) {
                const int tile = c >> lcpt, s0 = (c & ((1 << lcpt) - 1)) * len;
                double* sl = slot + (c - c0) * xb_slot(D);
                if (tile < noff) xb_offdiag_chunk<D, PBC>(N, c2, x1, Shi, Tb, rec, A, A + 1 + tile, s0, len, sl);
                else xb_diag_force_chunk<D, PBC>(N, x1, rec, dbuf, A + 1, s0, len, sl);
            }
            __syncthreads();
            if (cend < nchunk || A < 0) {   // not the round's last sub-round before a solve: everything now
                xb_reduce<D>(N, A, noff, dforce, lcpt, c0, cend, slot, acc, hsum, 0, -1, tid, nthr);
                __syncthreads();
            } else {
                break;   // c0 stays at the last sub-round, reduced below
            }
        }
        if (A < 0) {
            flush_rows(0);
            break;
        }
        // h of the last sub-round (one wave) beside the diagonal
        // coefficients; then the solve on wave 0 while the other waves add
        // that sub-round's F_1 and F_P partials
        const bool pending = c0 < nchunk;
        const int cend = min(nchunk, c0 + nslot);
        if (pending && wave == nw - 1)
            xb_reduce<D>(N, A, noff, dforce, lcpt, c0, cend, slot, acc, hsum, 0, 1, lane, WAVE);
        xb_diag_coefs<D, PBC>(N, c2, x1, Shi, Tb, rec, A, dbuf);
        __syncthreads();
        if (wave == 0) {
            xb_solve<D>(N, A, hsum, dbuf, rec, out);
            if (nw == 1 && pending) xb_reduce<D>(N, A, noff, dforce, lcpt, c0, cend, slot, acc, hsum, 1, -1, lane, WAVE);
        } else if (pending) {
            xb_reduce<D>(N, A, noff, dforce, lcpt, c0, cend, slot, acc, hsum, 1, -1, tid - WAVE, nthr - WAVE);
        }
        __syncthreads();
        if (tid < REC_B) hsum[tid] = 0.0;
        if (A + 1 < T) flush_rows(A + 1);   // (its diagonal forces were added this round)
    }
    // ---- bead P: the links P[l][l+1] = 1 - G_{l+1} (record l) and the sums
    // into f, f_P += k L (S + links)
    for (int l = tid; l < N; l += nthr) {
        double sP[D];
#pragma unroll
        for (int ax = 0; ax < D; ++ax) sP[ax] = acc[2 * D * REC_B + ax * N + l];
        if (l + 1 < N) {
            const double* rp = rec + l * REC;
            const double pP = 1.0 - rp[D + 2];
#pragma unroll
            for (int ax = 0; ax < D; ++ax) sP[ax] = fma(pP, xb_image<PBC>(x1[ax * N + l + 1] - rp[ax]), sP[ax]);
        }
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            const size_t i = (size_t)ax * N + l;
            fP[i] = fma(sP[ax], kL, fP[i]);
        }
    }
    __syncthreads();
}

}  // namespace pimdb
