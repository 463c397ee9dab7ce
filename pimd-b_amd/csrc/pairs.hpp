// pairs.hpp -- pair forces of one time slice on one wave, Newton's third law,
// in a fixed order.
//
// Reference: Simulation::updatePhysicalForces (src/simulation.cpp:428-455):
// every unordered pair (i<j) within the cutoff adds -gradV(r_i - r_j) to i and
// its negative to j.
//
// Particles are cut into tiles of 64.  For a tile pair (I, J), I <= J, lane l
// owns row particle I*64+l (position and force in registers) and holds one
// column particle of tile J, whose position, force accumulator and index move
// one lane along the wave after every step (a DPP wave rotation), so at every
// step the 64 lanes meet 64 distinct column particles and no two lanes touch
// the same accumulator.  An off-diagonal tile pair takes 64 steps (all 64 x 64
// pairs); a diagonal tile takes offsets 1..32 (offset 32 on lanes < 32 only),
// each unordered pair of the tile exactly once.  Which way the rotation turns
// does not matter: the column index travels with the data.
#pragma once

#include <type_traits>

#include "device_common.hpp"

namespace pimdb {

constexpr int PTILE = 64;

struct PairParams {
    double L, invL;
    double cutoff;
    double cut2;                       // r^2 < cut2  <=>  sqrt(r^2) < cutoff (cutoff_threshold)
    double int_k, strength;
    int pbc, guard;                    // guard: exact image choice near L/2 (PBC without a cutoff)
};

__host__ __device__ inline PairParams make_pair_params(double L, int pbc, double cutoff, double cut2, double int_k,
                                                       double strength) {
    PairParams pp;
    pp.L = L;
    pp.invL = 1.0 / L;
    pp.pbc = pbc;
    pp.cutoff = cutoff;
    pp.cut2 = cut2;
    pp.int_k = int_k;
    pp.strength = strength;
    pp.guard = (pbc && cutoff < 0.0) ? 1 : 0;
    return pp;
}

// Minimum image.  With a cutoff <= L/2 a displacement whose image choice is
// ambiguous (|dx| within rounding of L/2) lies outside the cutoff whichever
// image is taken, so there dx - L rint(dx/L) (one rounding less than the
// reference's dx - L floor(dx/L + 1/2)) serves; without a cutoff the guarded
// exact path keeps the reference's image.
// IMG: 0 no PBC, 1 PBC with a cutoff, 2 PBC without a cutoff (pair_image()).
template <int IMG>
__device__ __forceinline__ double min_image_pair(double dx, const PairParams& pp) {
    if constexpr (IMG == 0) return dx;
    else if constexpr (IMG == 2) return min_image_fast(dx, pp.L, pp.invL);
    else return fma(-pp.L, rint(dx * pp.invL), dx);
}

__host__ __device__ inline int pair_image(const PairParams& pp) { return !pp.pbc ? 0 : (pp.guard ? 2 : 1); }

// 1/sqrt(r2) to double precision (to 1 ulp; the pair forces use the cheaper
// rsqrt_pair below): the hardware estimate y0 and one third-order correction.  With e = 1 - r2 y0^2, 1/sqrt(r2) = y0 (1-e)^(-1/2)
// = y0 (1 + e/2 + 3e^2/8 + 5e^3/16 + ...); the dropped term is 5e^3/16, far
// below an ulp for the estimate's |e| ~ 2^-23.
__device__ __forceinline__ double rsqrt_refined(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double e = fma(-r2, y0 * y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// 1/sqrt(r2) for the pair forces: the estimate and one Newton step,
// y0 (1 + e/2); the dropped 3e^2/8 is about 5e-15 relative for |e| ~ 2^-23,
// far inside the forces' 1e-10.  One instruction fewer than rsqrt_refined.
__device__ __forceinline__ double rsqrt_pair(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double e = fma(-r2, y0 * y0, 1.0);
    return fma(0.5 * y0, e, y0);
}

// The reference's strict test |d| < cutoff (simulation.cpp:444) on a squared
// distance r2: the square root is correctly rounded and increasing, so
// sqrt(r2) < cutoff exactly when r2 is below the smallest double whose root
// reaches the cutoff.  That threshold is found once on the host
// (cutoff_threshold, ctx_common.hpp); the pair loop compares r2 with it.

// gradV(r_i - r_j) = g d for the pair (i, j): d = r_i - r_j (min-imaged) and
// g, if the pair is inside the cutoff (distance < cutoff, or cutoff < 0);
// false otherwise.  The force on i is -g d, on j +g d.
template <int D, int KIND, int IMG>
__device__ __forceinline__ bool pair_grad(const double (&xi)[D], const double (&xj)[D], const PairParams& pp,
                                          double (&d)[D], double& g) {
#pragma unroll
    for (int a = 0; a < D; ++a) d[a] = min_image_pair<IMG>(xi[a] - xj[a], pp);
    const double r2 = norm2_fma<D>(d);
    if (!(pp.cutoff < 0.0 || r2 < pp.cut2)) return false;
    const double inv = rsqrt_pair(r2);
    g = pair_grad_factor_fast<KIND>(r2 * inv, inv, pp.int_k, pp.strength);
    return true;
}

// ---- the pair gradient in box-fractional coordinates (PBC with a cutoff) ----
// With positions s = x / L the minimum image is t - rint(t), t = s_i - s_j
// (a subtraction and a rounding per axis instead of four operations), and
// every power of L folds into per-box constants: gradV(d) = g d = g_s d_s
// with d_s = d / L and g_s = g L.  With y = L / r (one reciprocal square root
// of |d_s|^2), q = y^2 and r_s = r / L:
//   harmonic  g_s = k_pair L
//   dipole    g_s = (-3 strength / L^4) q^2 y
//   Aziz      g_s = y T1 + U                       (r >= D rm)
//             g_s = y T1 + (U - t y q^4 pc(q)) F   (r < D rm; U alone outside
//                                                   the reference's hard-core
//                                                   range r/rm in (1e-7, 0.01))
// where, with ER = epsilon / rm and a = rm / L (pair_grad_factor_fast's
// expression rewritten in y):
//   T1 = -A alpha ER exp(-alpha r / rm)              (the repulsion)
//   U  = q^4 (b0 + b1 q + b2 q^2),  b = ER a^7 (6 C6, 8 C8 a^2, 10 C10 a^4)
//   pc = 2 D ER a^8 (C6 + C8 a^2 q + C10 a^4 q^2)
//      = D a (b0 / 3 + b1 q / 4 + b2 q^2 / 5)
//   t  = D a y - 1,  F = exp(-t^2)                   (aziz.h:25-34)
// The damping-range terms are formed from b and D a inside that branch, so
// the loop holds six per-box constants.
// Both exponentials are 2^(k + f), k = rint(a), f = a - k exactly, with a the
// exponent in base 2 and 2^f from frac_exp2 below.  The cutoff test compares
// |d_s|^2 with (cutoff threshold) / L^2.
struct FracParams {
    double cut2;            // squared cutoff threshold over L^2
    double kh, kd;          // harmonic and dipole factors
    double ex;              // -alpha L / rm * log2(e): the repulsion's exponent per unit r_s
    double b0, b1, b2;      // undamped dispersion
    double b1v, b2v;        // b1 and b2 again, for copies in vector registers (vgpr_copy)
    double dl;              // D rm / L: the damping range in r_s, and t = dl y - 1
};

__host__ __device__ inline FracParams make_frac_params(const PairParams& pp) {
    FracParams f;
    const double L = pp.L, iL = pp.invL;
    f.cut2 = pp.cut2 * iL * iL;
    f.kh = pp.int_k * L;
    f.kd = -3.0 * pp.strength * (iL * iL) * (iL * iL);
    const double ER = Aziz::epsilon / Aziz::rm;
    const double a = Aziz::rm * iL, a2 = a * a, a7 = a * a2 * a2 * a2;
    f.ex = -Aziz::alpha * (L / Aziz::rm) * 1.4426950408889634;
    f.b0 = 6.0 * Aziz::C6 * ER * a7;
    f.b1 = 8.0 * Aziz::C8 * ER * a7 * a2;
    f.b1v = f.b1;
    f.b2 = 10.0 * Aziz::C10 * ER * a7 * a2 * a2;
    f.b2v = f.b2;
    f.dl = Aziz::D * a;
    return f;
}

// The Aziz repulsion's prefactor -A alpha epsilon / rm, folded into frac_exp2.
constexpr double AZ_T1C = -Aziz::A * Aziz::alpha * (Aziz::epsilon / Aziz::rm);
constexpr double LN2D = 6.931471805599453e-01;

// AZ_T1C 2^f for |f| <= 1/2: exp_pair's degree-9 minimax polynomial for e^r
// on |r| <= ln(2)/2 taken at r = f ln 2 (coefficient j times (ln 2)^j) and
// scaled by AZ_T1C, so the same relative accuracy (1.4e-14, plus a few ulp
// from the folded constants).
__device__ __forceinline__ double frac_exp2(double f) {
    constexpr double l1 = LN2D, l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1, l6 = l5 * l1,
                     l7 = l6 * l1, l8 = l7 * l1, l9 = l8 * l1;
    double p = fma_sconst(AZ_T1C * 2.7482187570680464e-06 * l9, f, AZ_T1C * 2.488401002638418e-05 * l8);
    p = fma_sconst(p, f, AZ_T1C * 1.9841596395525616e-04 * l7);
    p = fma_sconst(p, f, AZ_T1C * 1.3888802828792633e-03 * l6);
    p = fma_sconst(p, f, AZ_T1C * 8.333333002823904e-03 * l5);
    p = fma_sconst(p, f, AZ_T1C * 4.166666703246359e-02 * l4);
    p = fma_sconst(p, f, AZ_T1C * 1.6666666667826951e-01 * l3);
    p = fma_sconst(p, f, AZ_T1C * 4.999999999945797e-01 * l2);
    p = fma_sconst(p, f, AZ_T1C * 9.999999999998928e-01 * l1);
    p = fma_sconst(p, f, AZ_T1C * 1.0000000000000127);
    return p;
}

// gradV = g d for the pair (i, j) in fractional coordinates (xi, xj = x / L):
// d = min-imaged s_i - s_j and g = g_s as above when the pair is inside the
// cutoff (the return value), g = 0 otherwise.  Branch-free but for the
// damping range: outside the cutoff r > L/2 and every term is finite, and in
// the hard-core range the damped terms are finite with F = 0.
template <int D, int KIND>
__device__ __forceinline__ bool pair_grad_frac(const double (&xi)[D], const double (&xj)[D], const FracParams& fp,
                                               double (&d)[D], double& g) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double t = xi[a] - xj[a];
        d[a] = t - rint(t);
    }
    const double s = norm2_fma<D>(d);
    const bool in = s < fp.cut2;
    if constexpr (KIND == POT_HARMONIC) {
        g = fp.kh;
    } else if constexpr (KIND == POT_DIPOLE) {
        const double y = rsqrt_pair(s), q = y * y;
        g = fp.kd * (q * q * y);
    } else if constexpr (KIND == POT_AZIZ) {
        const double y = rsqrt_pair(s);
        const double rs = s * y;
        const double e1 = fp.ex * rs, k1 = rint(e1);
        const double T1 = ldexp(frac_exp2(e1 - k1), (int)k1);
        const double q = y * y, q2 = q * q, q4 = q2 * q2;
        double att = q4 * fma(fma(fp.b2v, q, fp.b1v), q, fp.b0);
        if (rs < fp.dl) {
            // The reference's hard-core range r / rm in (1e-7, 0.01), where it
            // keeps the repulsion alone, needs no test of its own: there
            // t = D rm / r - 1 > 123 and F = 0 exactly (F > 0 needs t < 27.4;
            // the exponent takes t clamped to 38, e2 >= -2084, which keeps
            // the conversion to int in range and changes no nonzero F),
            // while the other factors stay finite (y^8 < 1.2e63), so the
            // damped term is +-0 and g = y T1 to the bit, as with att = 0.
            const double t = fma(fp.dl, y, -1.0), tc = fmin(t, 38.0);
            const double e2 = (tc * -1.4426950408889634) * tc, k2 = rint(e2);
            const double F = ldexp(frac_exp2(e2 - k2), (int)k2) * (1.0 / AZ_T1C);
            const double pc = fma(fma(fp.b2 * 0.2, q, fp.b1 * 0.25), q, fp.b0 * (1.0 / 3.0)) * fp.dl;
            att = fma(-(t * y) * q4, pc, att) * F;
        }
        g = fma(y, T1, att);
    } else {
        g = 0.0;
    }
    g = in ? g : 0.0;
    return in;
}

// LDS pointer type: accesses through it are ds_ instructions that the
// compiler knows cannot alias global memory.
typedef __attribute__((address_space(3))) double lds_double;
// Global-memory pointer type: loads through it are global_load (counted by
// vmcnt only), not flat loads, whose lgkmcnt share would make the pair loop
// wait for its own LDS force updates on every step.
typedef const __attribute__((address_space(1))) double* gbl_cdptr;
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) unsigned lds_uint;

// All pairs of one slice on one wave.  xs: slice positions [D][N]; fo: the
// slice's force array [D][N], owned by this wave, overwritten with the pair
// forces.  Must be called by all 64 lanes together.
template <int D, int KIND, int IMG>
__device__ __forceinline__ void slice_pair_forces(const double* xs, double* fo, int N, const PairParams& pp, int lane) {
    for (int e = lane; e < D * N; e += WAVE) fo[e] = 0.0;
    __threadfence_block();
    const int T = (N + PTILE - 1) / PTILE;
    for (int I = 0; I < T; ++I) {
        const int i = I * PTILE + lane;
        const bool row_ok = i < N;
        double xi[D], racc[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = row_ok ? xs[a * N + i] : 0.0;
            racc[a] = 0.0;
        }
        for (int J = I; J < T; ++J) {
            const bool diag = J == I;
            const int s0 = diag ? 1 : 0, s1 = diag ? 33 : PTILE;
            int jc = J * PTILE + ((lane + s0) & (PTILE - 1));
            double xj[D], cacc[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                xj[a] = jc < N ? xs[a * N + jc] : 0.0;
                cacc[a] = 0.0;
            }
            for (int s = s0; s < s1; ++s) {
                if (row_ok && jc < N && (!diag || s < 32 || lane < 32)) {
                    double d[D], g;
                    if (pair_grad<D, KIND, IMG>(xi, xj, pp, d, g)) {
#pragma unroll
                        for (int a = 0; a < D; ++a) {
                            racc[a] = fma(-g, d[a], racc[a]);
                            cacc[a] = fma(g, d[a], cacc[a]);
                        }
                    }
                }
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    xj[a] = lane_rotate(xj[a]);
                    cacc[a] = lane_rotate(cacc[a]);
                }
                jc = lane_rotate(jc);
            }
            if (jc < N) {
#pragma unroll
                for (int a = 0; a < D; ++a) fo[a * N + jc] += cacc[a];
            }
            __threadfence_block();
        }
        if (row_ok) {
#pragma unroll
            for (int a = 0; a < D; ++a) fo[a * N + i] += racc[a];
        }
        __threadfence_block();
    }
}

// The pair gradient of the LDS pair walk: pair_grad on the positions as
// stored or, with periodic boundaries and a cutoff (IMG 1), pair_grad_frac on
// the positions scaled by 1/L as they are loaded.  Both give gradV = g d with
// g d in force units.
template <int D, int KIND, int IMG>
struct WalkGrad {
    PairParams pp;
    FracParams fp;
    __device__ __forceinline__ double load(const gbl_cdptr xs, int idx) const {
        if constexpr (IMG == 1) return xs[idx] * pp.invL;
        else return xs[idx];
    }
    __device__ __forceinline__ bool operator()(const double (&xi)[D], const double (&xj)[D], double (&d)[D],
                                               double& g) const {
        if constexpr (IMG == 1) return pair_grad_frac<D, KIND>(xi, xj, fp, d, g);
        else return pair_grad<D, KIND, IMG>(xi, xj, pp, d, g);
    }
};

// Doubles per tile position in the LDS force arrays of the slice walk: D
// rounded up to odd, so that 32 consecutive positions (64-bit LDS accesses
// of half a wave) cover 32 distinct bank pairs.
__host__ __device__ constexpr int pair_fstride(int D) { return D | 1; }

// All pairs of one slice on one wave, with the forces accumulated in the
// wave's own LDS array F (zeroed here; the pair forces are left in it: the
// caller adds the rest of the force and writes the slice once).  Used when
// every wave's slice fits the LDS.
//
// The tiles are cut from the particles in the order perm[0..N) (a
// permutation of 0..N-1, in LDS): tile position k holds particle perm[k].
// Positions are read by the particle itself, so any permutation gives the
// same set of pairs; forces are accumulated by tile position (F[k S + a],
// S = pair_fstride(D)), so the 64 lanes of a tile update 64 consecutive
// positions, and with S odd any 32 consecutive lanes fall on distinct LDS
// bank pairs, for every rotation offset.  The caller reads particle i's force
// at its position iperm[i], and orders the particles spatially
// (pair_order_key) so that the close pairs, where a branch of the pair force
// is taken, meet at few offsets of the rotation.
//
// Every pass holds two rows per lane against one rotating column, so that the
// column's DPP rotation and its LDS update (the two pairs' terms added first)
// serve two pairs:
//  - a diagonal tile I: lane l holds rows l and l+32 (mod 64) of the tile and
//    the column l+s.  Offsets s = 1..15 give the pairs at distance s (first
//    row) and 32-s (second row) along the tile's ring of 64 positions; s = 16
//    the distance 16 once (first row), and s = 0 the distance 32 (second row,
//    lanes < 32), each pair of the tile exactly once;
//  - off-diagonal tile pairs, two row tiles against one column tile: column
//    J >= 1 takes rows 0..J-1 for J even and rows 1..J-1 for J odd, in pairs
//    of consecutive tiles, and column 0 the odd row tiles (1,3), (5,7), ...
//    (the pairs {0, J}, J odd, oriented the other way), so that at most one
//    pass (column 0's last, when the number of odd tiles is odd) has a single
//    row.
template <int D, int KIND, int IMG>
struct SliceWalk {
    gbl_cdptr xs;
    int N;
    const WalkGrad<D, KIND, IMG>& wg;
    lds_double* F;
    const lds_int* perm;
    int lane;
    double xj[D];   // the lane's column: positions,
    unsigned jd;    // and the LDS byte address of its first element in F
                    // (F + kc S, so the updates need no address arithmetic);
                    // at or past F + N S: none

    __device__ __forceinline__ bool load_row(int k, double (&xr)[D]) const {
        const bool ok = k < N;
        const int i = ok ? perm[k] : 0;
#pragma unroll
        for (int a = 0; a < D; ++a) xr[a] = ok ? wg.load(xs, a * N + i) : 0.0;
        return ok;
    }
    __device__ __forceinline__ void flush_row(int k, bool ok, const double (&r)[D]) const {
        if (ok) {
#pragma unroll
            for (int a = 0; a < D; ++a)
                __hip_atomic_fetch_add(&F[k * pair_fstride(D) + a], r[a], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __device__ __forceinline__ static unsigned lds_addr(const lds_double* q) { return (unsigned)(size_t)q; }
    __device__ __forceinline__ void load_col(int kc) {
        const int jc = kc < N ? perm[kc] : N;
        jd = lds_addr(F + kc * pair_fstride(D));
#pragma unroll
        for (int a = 0; a < D; ++a) xj[a] = jc < N ? wg.load(xs, a * N + jc) : 0.0;
    }
    __device__ __forceinline__ bool col_ok() const { return jd < lds_addr(F + N * pair_fstride(D)); }
    __device__ __forceinline__ void rotate() {
#pragma unroll
        for (int a = 0; a < D; ++a) xj[a] = lane_rotate(xj[a]);
        jd = (unsigned)lane_rotate((int)jd);
    }
    // the pair (row xi, column) if act
    __device__ __forceinline__ void one(const double (&xi)[D], double (&ri)[D], bool act) {
        if (act) {
            double d[D], g;
            if (wg(xi, xj, d, g)) {
                lds_double* Fj = (lds_double*)(size_t)jd;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    ri[a] = fma(-g, d[a], ri[a]);
                    __hip_atomic_fetch_add(&Fj[a], g * d[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    // the pairs (row xa, column) if acta and (row xb, column) if actb
    __device__ __forceinline__ void two(const double (&xa)[D], const double (&xb)[D], double (&ra)[D],
                                        double (&rb)[D], bool acta, bool actb) {
        if (acta || actb) {
            double da[D], db[D], ga = 0.0, gb = 0.0;
            const bool ina = wg(xa, xj, da, ga) && acta;
            const bool inb = wg(xb, xj, db, gb) && actb;
            if (ina || inb) {
                if (!ina) ga = 0.0;
                if (!inb) gb = 0.0;
                lds_double* Fj = (lds_double*)(size_t)jd;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    ra[a] = fma(-ga, da[a], ra[a]);
                    rb[a] = fma(-gb, db[a], rb[a]);
                    __hip_atomic_fetch_add(&Fj[a], fma(ga, da[a], gb * db[a]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }

    // the pairs of diagonal tile I
    __device__ __forceinline__ void diag_tile(int I) {
        const int ka = I * PTILE + lane, kb = I * PTILE + ((lane + PTILE / 2) & (PTILE - 1));
        double xa[D], xb[D], ra[D] = {}, rb[D] = {};
        const bool oka = load_row(ka, xa), okb = load_row(kb, xb);
        load_col(ka);
        one(xb, rb, okb && lane < PTILE / 2 && col_ok());   // s = 0: distance 32
        rotate();
        if ((I + 1) * PTILE <= N) {
            for (int s = 1; s < PTILE / 4; ++s) {
                two(xa, xb, ra, rb, true, true);
                rotate();
            }
            one(xa, ra, true);   // s = 16
        } else {
            for (int s = 1; s < PTILE / 4; ++s) {
                const bool c = col_ok();
                two(xa, xb, ra, rb, oka && c, okb && c);
                rotate();
            }
            one(xa, ra, oka && col_ok());
        }
        flush_row(ka, oka, ra);
        flush_row(kb, okb, rb);
    }
    // row tiles A and B (B < 0: none) against column tile J, all distinct
    __device__ __forceinline__ void off_tiles(int A, int B, int J) {
        const int ka = A * PTILE + lane, kb = B * PTILE + lane;
        double xa[D], xb[D], ra[D] = {}, rb[D] = {};
        const bool oka = load_row(ka, xa), okb = B >= 0 && load_row(kb, xb);
        load_col(J * PTILE + lane);
        const bool full = (A + 1) * PTILE <= N && (J + 1) * PTILE <= N && (B < 0 || (B + 1) * PTILE <= N);
        if (B < 0) {
            for (int s = 0; s < PTILE; ++s) {
                one(xa, ra, oka && col_ok());
                rotate();
            }
        } else if (full) {
            for (int s = 0; s < PTILE; ++s) {
                two(xa, xb, ra, rb, true, true);
                rotate();
            }
        } else {
            for (int s = 0; s < PTILE; ++s) {
                const bool c = col_ok();
                two(xa, xb, ra, rb, oka && c, okb && c);
                rotate();
            }
        }
        flush_row(ka, oka, ra);
        if (B >= 0) flush_row(kb, okb, rb);
    }
};

template <int D, int KIND, int IMG>
__device__ __forceinline__ void slice_pair_forces_lds(const double* xs_, int N, const WalkGrad<D, KIND, IMG>& wg,
                                                      int lane, lds_double* F, const lds_int* perm) {
    for (int e = lane; e < pair_fstride(D) * N; e += WAVE) F[e] = 0.0;
    __builtin_amdgcn_wave_barrier();
    SliceWalk<D, KIND, IMG> w{(gbl_cdptr)xs_, N, wg, F, perm, lane, {}, 0u};
    const int T = (N + PTILE - 1) / PTILE;
    for (int I = 0; I < T; ++I) w.diag_tile(I);
    for (int J = 1; J < T; ++J)
        for (int A = J & 1; A < J; A += 2) w.off_tiles(A, A + 1, J);
    for (int A = 1; A < T; A += 4) w.off_tiles(A, A + 2 < T ? A + 2 : -1, 0);
    __builtin_amdgcn_wave_barrier();
}

// Spatial order for the tiles of slice_pair_forces_lds: the key of particle
// i (position xi, box L) is its cell in a row-major grid of about one cell
// per particle, 4 cells along the first axis, and the particle index below
// it, so the keys are distinct.  Four cells along the first axis put a
// particle's neighbours at tile offsets +-1 and +-3 (that axis, periodic),
// +-4 (the second) and +-4 ny (the third): close pairs then meet at few
// rotation offsets, and tiles two or more apart along the last axis hold no
// close pairs at all.
template <int D>
__device__ __forceinline__ unsigned pair_order_key(const double (&xi)[D], double invL, int N, int i) {
    int ny = 1, nz = 1;
    if constexpr (D == 2) ny = max(1, (N + 3) / 4);
    if constexpr (D == 3) ny = nz = max(1, (int)sqrt((double)N / 4.0 + 0.5));
    const int n[3] = {D == 1 ? N : 4, ny, nz};
    unsigned cell = 0;
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
        double f = xi[a] * invL;
        f -= floor(f);
        const int c = min(n[a] - 1, (int)(f * n[a]));
        cell = cell * (unsigned)n[a] + (unsigned)c;
    }
    return (cell << 16) | (unsigned)i;
}

}  // namespace pimdb
